"""Graph-replayed steps (mhd2d_step_graph) and the loop test on the device.

- nsteps graph launches equal nsteps eager mhd2d_step calls bit for bit, in
  the plain form and in the mixed form with the learned screen, across more
  than one six-step rotation of the buffer roles and for chunked replays
- a step that starts at t >= tend changes nothing: graph chunks that run past
  tend end on the same state, t and step count as the eager loop that stops
  at tend, further passes leave the state untouched, and a later, larger tend
  continues from the right buffers
- the mixed form's double-double CT keeps div B from accumulating
- weights set after the capture take effect in the replays
- the stage's second stream gives the same bits as the one-stream step,
  also on a flow whose limiter lists are busy (the rotor)
- closing a context gives its device memory back
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make(pkg, prec, ind, N=48):
    s = pkg.Solver(N, N, flux_type=pkg.HLLD, M=50.0, CFL=0.2, pp=1, indicator=ind, precision=prec, Ck=1.0)
    if ind == pkg.IND_MLP:
        s.mlp_set_weights(pkg.load_indicator_weights())
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 7])
    s.rk3_prologue()
    return s


def same(a, b):
    ua, xa, ya = a.download()
    ub, xb, yb = b.download()
    return (np.array_equal(ua, ub) and np.array_equal(xa, xb) and np.array_equal(ya, yb)
            and a.get_time() == b.get_time())


@pytest.mark.parametrize("form", ["fp64", "mixed_nn"])
def test_graph_equals_eager(pkg, form):
    prec, ind = (pkg.FP64, pkg.IND_KXRCF) if form == "fp64" else (pkg.MIXED, pkg.IND_MLP)
    a, b = make(pkg, prec, ind), make(pkg, prec, ind)
    for _ in range(15):
        a.step()
    b.step_graph(0)                                # capture the six phases ahead, run nothing
    assert b.get_time()[2] == 0
    for k in (5, 1, 7, 2):
        b.step_graph(k)
    print(f"[graph] {form}: after 15 steps t = {a.get_time()[0]:.6f}")
    assert a.get_time()[2] == 15
    assert same(a, b)


def test_steps_past_tend_change_nothing(pkg):
    a, b = make(pkg, pkg.FP64, pkg.IND_KXRCF), make(pkg, pkg.FP64, pkg.IND_KXRCF)
    tend = 0.05
    n = 0
    t = 0.0
    while t < tend:
        t = a.step(tend, sync=True)
        n += 1
    tb = b.step_graph(n + 9, tend, sync=True)     # 9 passes start at t = tend
    print(f"[graph] tend {tend}: {n} steps, t = {t}, graph t = {tb}")
    assert t == tend and tb == tend and b.get_time()[2] == n
    assert same(a, b)
    u0 = a.download()
    for _ in range(4):                             # eager passes past tend
        a.step(tend)
    assert a.get_time()[2] == n
    assert all(np.array_equal(x, y) for x, y in zip(u0, a.download()))
    assert same(a, b)
    # a larger tend continues from the right buffers, eager and graph alike
    for _ in range(7):
        a.step(0.08)
    b.step_graph(7, 0.08)
    assert same(a, b)


def test_mixed_divergence_does_not_accumulate(pkg):
    """the plain form's div B is a random walk of the face roundings (as in
    the reference: it grows as the square root of the step count); the mixed
    form's double-double CT update keeps it at one rounding"""
    out = {}
    for name, prec in (("fp64", pkg.FP64), ("mixed", pkg.MIXED)):
        s = make(pkg, prec, pkg.IND_KXRCF, N=128)
        div = []
        for k in (10, 90, 300):
            s.step_graph(k)
            div.append(s.diagnostics()["totaldiv"])
        out[name] = div
        s.close()
    print(f"[graph] div B after 10, 100, 400 steps: fp64 {out['fp64']}, mixed {out['mixed']}")
    assert out["fp64"][2] > 3 * out["fp64"][0]
    assert out["mixed"][2] < 2 * out["mixed"][0]
    assert out["mixed"][2] < 0.25 * out["fp64"][2]


def test_graph_sees_new_weights(pkg):
    """the captured graphs read the indicator weights through the context's
    weight buffer: weights set after the capture take effect in the replays"""
    a, b, c = (make(pkg, pkg.MIXED, pkg.IND_MLP) for _ in range(3))
    b.step_graph(0)
    other = pkg.load_indicator_weights().copy()
    other[-1] += 50.0                   # output bias: the screen passes most cells
    for s in (a, b):
        s.mlp_set_weights(other)
    for _ in range(4):
        a.step()
        c.step()                        # the trained weights
    b.step_graph(4)
    assert same(a, b)
    pa, pb, pc = a.mlp_probability(), b.mlp_probability(), c.mlp_probability()
    assert np.array_equal(pa, pb)
    assert (pa > 0.5).mean() > 0.5 > (pc > 0.5).mean()


@pytest.mark.parametrize("form", ["fp64", "mixed_nn", "mixed_kxrcf"])
def test_two_streams_equal_one_stream(pkg, form):
    """the stage's second stream (vertex EMFs, CT, detection) against the
    same step run on one stream (the per-kernel timing mode): bit-identical,
    eager and graph-replayed"""
    prec, ind = {"fp64": (pkg.FP64, pkg.IND_KXRCF), "mixed_nn": (pkg.MIXED, pkg.IND_MLP),
                 "mixed_kxrcf": (pkg.MIXED, pkg.IND_KXRCF)}[form]
    a, b, c = (make(pkg, prec, ind) for _ in range(3))
    a.profile(True)                     # one stream
    for _ in range(7):
        a.step()
        b.step()
    c.step_graph(7)
    a.profile(False)
    assert same(a, b) and same(a, c)


@pytest.mark.parametrize("form", ["mixed_kxrcf", "mixed_nn", "fp64"])
def test_two_streams_equal_one_stream_rotor(pkg, form):
    """the same on the reference's rotor (init4.txt, positivity, TVB M = 10
    gated by KXRCF or the screen), where TVB changes cells at every stage:
    the mixed form's list kernels limit those cells in place after the
    detection has read the stage value, and the next stage combines their
    unlimited value from the companion buffer; the two-stream order, the
    one-stream order and the graph replays must give the same bits"""
    import importlib
    prec, ind = {"fp64": (pkg.FP64, pkg.IND_KXRCF), "mixed_nn": (pkg.MIXED, pkg.IND_MLP),
                 "mixed_kxrcf": (pkg.MIXED, pkg.IND_KXRCF)}[form]
    d = pkg.ic_defaults(pkg.IC_ROTOR)
    W = pkg.load_indicator_weights()
    z = np.load(pkg.WEIGHTS_PATH, allow_pickle=False)

    def mk():
        s = pkg.Solver(64, 64, flux_type=pkg.HLLD, bc=d["bc"], domain=d["domain"], M=10.0, CFL=0.15, pp=1,
                       indicator=ind, precision=prec, Ck=1.0, mlp_band=(float(z["mlp_lo"]), float(z["mlp_hi"])))
        s.mlp_set_weights(W)
        s.init_data(pkg.IC_ROTOR)
        s.rk3_prologue()
        return s
    a, b, c = mk(), mk(), mk()
    a.profile(True)
    for _ in range(12):
        a.step()
        b.step()
    c.step_graph(12)
    a.profile(False)
    assert same(a, b) and same(a, c)
    st = importlib.import_module("freljord_cfd_arsenal_amd.indicator").limiter_input_stats(a)
    print(f"[rotor] {form}: after 12 steps KXRCF flags {st['kxrcf']}, TVB changes {st['label']}")
    assert st["label"] > 0          # the flow keeps the limiter lists busy


def test_contexts_free_their_memory(pkg):
    """closing a context after graph-replayed and two-stream eager steps
    gives its device memory back (graphs are captured on one stream: a graph
    with parallel branches kept its buffers from being released)"""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    base = None
    for _ in range(3):
        s = make(pkg, pkg.MIXED, pkg.IND_MLP, N=512)
        s.step_graph(0)
        s.step_graph(3)
        s.step()
        s.sync()
        s.close()
        f = free_bytes()
        base = f if base is None else base
        assert f >= base - (64 << 20)
