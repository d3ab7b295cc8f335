"""Initial-condition library, vector-potential init, L2-error diagnostic and
the binary checkpoint on the GPU, against the reference fixtures.

- init_data with init5-9b.txt (P0 data where shock = 1): relative L2 1e-13
  of the reference's projection (the device evaluates sqrt and the
  transcendental functions with its own libm)
- calculate_Az (magnetic loop) and calculate_L2error (vortex): 1e-13 / 1e-12
- checkpoint: save, restart in a fresh context and continue is bit-identical
  to the uninterrupted run (plain and mixed forms); the trailer is zlib's
  CRC-32 of header + payload; a corrupted or mismatched file is rejected
  before the context is touched
"""
import struct
import zlib

import numpy as np
import pytest

from conftest import golden, interior, rel_l2, unpack

pytestmark = pytest.mark.gpu

IC_LIBRARY = [("blast", 5), ("cloud", 6), ("cloud2", 61), ("loop", 7), ("briowu", 8), ("briowu_y", 81),
              ("ryujones", 9), ("ryujones_y", 91)]


@pytest.mark.parametrize("tag,ic", IC_LIBRARY)
def test_init_library(pkg, tag, ic):
    g = golden(f"init_{tag}_16")
    d = pkg.ic_defaults(ic)
    s = pkg.Solver(16, 16, bc=d["bc"], domain=d["domain"], M=d["M"])
    s.init_data(ic)
    uh, Bx, By = s.download()
    ref = unpack(g["uh"])
    assert rel_l2(interior(uh), interior(ref)) < 1e-13
    assert rel_l2(Bx, g["Bx"]) < 1e-13 and rel_l2(By, g["By"]) < 1e-13
    if d["shock"]:
        assert np.abs(interior(uh)[:, :, 1:, :]).max() == 0.0


def test_calculate_az(pkg):
    g = golden("az_loop_16")
    s = pkg.Solver(16, 16, domain=tuple(g["domain"]))
    s.upload(unpack(g["uh"]), g["Bx0"], g["By0"])
    s.calculate_Az()
    _, Bx, By = s.download()
    assert rel_l2(Bx, g["Bx"]) < 1e-13 and rel_l2(By, g["By"]) < 1e-13
    # modes 2-3 are left as they were
    assert np.array_equal(Bx[:, :, 1:], g["Bx0"][:, :, 1:]) and np.array_equal(By[:, :, 1:], g["By0"][:, :, 1:])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_calculate_l2error(pkg, prec):
    g = golden("l2err_vortex_16")
    s = pkg.Solver(16, 16, domain=tuple(g["domain"]), precision=pkg.FP64 if prec == "fp64" else pkg.MIXED)
    s.upload(unpack(g["uh"]), g["Bx"], g["By"])
    tol = 1e-12 if prec == "fp64" else 1e-6
    for t20, ref in zip(g["t20"], g["L2"]):
        got = s.calculate_L2error(float(t20))
        assert np.all(np.abs(got - ref) <= tol * ref + 1e-14), (t20, got, ref)


def _run(pkg, prec, ind, steps, pre=None):
    s = pkg.Solver(32, 32, flux_type=pkg.HLLD, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi), M=50.0, CFL=0.2, pp=1,
                   precision=prec, indicator=ind, Ck=1.0)
    if pre is None:
        s.init_data(pkg.IC_OT_PERTURBED, [5e-2, 77])
        s.rk3_prologue()
    else:
        s.load_checkpoint(pre)
    for _ in range(steps):
        s.step()
    return s


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_checkpoint_restart_bit_exact(pkg, tmp_path, prec):
    P = pkg.FP64 if prec == "fp64" else pkg.MIXED
    path = str(tmp_path / "ck.bin")
    a = _run(pkg, P, pkg.IND_KXRCF, 4)
    a.save_checkpoint(path)
    for _ in range(4):
        a.step()
    b = _run(pkg, P, pkg.IND_KXRCF, 4, pre=path)
    ua, xa, ya = a.download()
    ub, xb, yb = b.download()
    assert np.array_equal(ua, ub) and np.array_equal(xa, xb) and np.array_equal(ya, yb)
    ta, tb = a.get_time(), b.get_time()
    assert ta[0] == tb[0] and ta[2] == tb[2] == 8


def test_checkpoint_format(pkg, tmp_path):
    path = tmp_path / "ck.bin"
    s = _run(pkg, pkg.FP64, pkg.IND_KXRCF, 2)
    s.save_checkpoint(str(path))
    raw = path.read_bytes()
    N = 32
    W = H = N + 2
    pay = (52 * W * H + 2 * 3 * (N + 1) * N) * 8
    assert len(raw) == 128 + pay + 4
    assert raw[:8] == b"MHD2DCK1"
    version, hb, Nx, Ny, deg, npl = struct.unpack_from("<IIiiii", raw, 8)
    assert (version, hb, Nx, Ny, deg, npl) == (1, 128, N, N, 2, 52)
    t, dt = struct.unpack_from("<dd", raw, 88)
    count, nbytes = struct.unpack_from("<qQ", raw, 104)
    assert (count, nbytes) == (2, pay) and t == s.get_time()[0]
    assert struct.unpack_from("<I", raw, 128 + pay)[0] == zlib.crc32(raw[:128 + pay])
    # the payload is the downloaded state in the documented order
    uh, Bx, By = s.download()
    planes = np.frombuffer(raw, dtype="<f8", count=52 * W * H, offset=128).reshape(52, H, W)
    order = [(d, n) for n in range(8) for d in range(6)] + [(6, 5), (8, 5), (7, 6), (9, 6)]
    for k, (d, n) in enumerate(order):
        assert np.array_equal(planes[k].T, uh[:, :, d, n])
    fx = np.frombuffer(raw, dtype="<f8", count=3 * (N + 1) * N, offset=128 + 52 * W * H * 8)
    assert np.array_equal(fx.reshape(Bx.shape, order="F"), Bx)
    # corruption and mismatches are rejected and leave the context untouched
    bad = bytearray(raw)
    bad[128 + 1000] ^= 0x10
    (tmp_path / "bad.bin").write_bytes(bytes(bad))
    before = s.download()[0]
    with pytest.raises(pkg.MHD2DError, match="CRC"):
        s.load_checkpoint(str(tmp_path / "bad.bin"))
    (tmp_path / "short.bin").write_bytes(raw[:5000])
    with pytest.raises(pkg.MHD2DError):
        s.load_checkpoint(str(tmp_path / "short.bin"))
    assert np.array_equal(s.download()[0], before)
    other = pkg.Solver(16, 16)
    with pytest.raises(pkg.MHD2DError, match="grid"):
        other.load_checkpoint(str(path))
    # a context with another domain, boundaries or Riemann solver rejects it
    dom = (0.0, 2 * np.pi, 0.0, 2 * np.pi)
    for kw, why in (({"domain": (0.0, 1.0, 0.0, 1.0)}, "domain"),
                    ({"domain": dom, "bc": (2, 2, 2, 2)}, "boundary"),
                    ({"domain": dom, "flux_type": pkg.HLL}, "flux_type")):
        args = dict(flux_type=pkg.HLLD, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_KXRCF, Ck=1.0)
        args.update(kw)
        o = pkg.Solver(N, N, **args)
        with pytest.raises(pkg.MHD2DError, match=why):
            o.load_checkpoint(str(path))


def test_checkpoint_of_a_flagged_run(pkg, tmp_path):
    """a numerical flag (a positive status of the time read) does not stop
    the save: the file is written and restarts bit for bit, flag included
    in nothing but the status"""
    s = _run(pkg, pkg.FP64, pkg.IND_KXRCF, 2)
    uh, Bx, By = s.download()
    neg = uh.copy()
    neg[7, 7, 0, 4] = 0.1 * neg[7, 7, 0, 4] - 5.0   # negative pressure at this cell
    s.upload(neg, Bx, By)
    assert s.diagnostics()["flags"] & pkg.FLAG_NEG_PRESSURE
    path = tmp_path / "flagged.bin"
    s.save_checkpoint(str(path))
    assert path.exists() and path.stat().st_size > 128
    r = pkg.Solver(32, 32, flux_type=pkg.HLLD, domain=(0.0, 2 * np.pi, 0.0, 2 * np.pi), M=50.0, CFL=0.2, pp=1,
                   indicator=pkg.IND_KXRCF, Ck=1.0)
    r.load_checkpoint(str(path))
    ur, xr, yr = r.download()
    us, xs, ys = s.download()
    assert np.array_equal(ur, us) and np.array_equal(xr, xs) and np.array_equal(yr, ys)
    assert r.get_time()[2] == s.get_time()[2] == 2


def test_checkpoint_version2_mixed(pkg, tmp_path):
    """a mixed-form context writes version 2: the v1 payload, then the low
    parts of the double-double mode-1 face values, each at most half an ulp
    of its high part; a plain context reads the file and ignores them"""
    path = tmp_path / "ck2.bin"
    s = _run(pkg, pkg.MIXED, pkg.IND_KXRCF, 3)
    s.save_checkpoint(str(path))
    raw = path.read_bytes()
    N = 32
    W = H = N + 2
    n1 = 52 * W * H + 2 * 3 * (N + 1) * N
    pay = (n1 + 2 * (N + 1) * N) * 8
    assert len(raw) == 128 + pay + 4
    assert struct.unpack_from("<I", raw, 8)[0] == 2
    assert struct.unpack_from("<Q", raw, 112)[0] == pay
    assert struct.unpack_from("<I", raw, 128 + pay)[0] == zlib.crc32(raw[:128 + pay])
    _, Bx, By = s.download()
    lo = np.frombuffer(raw, dtype="<f8", count=2 * (N + 1) * N, offset=128 + n1 * 8)
    lx = lo[:(N + 1) * N].reshape((N + 1, N), order="F")
    ly = lo[(N + 1) * N:].reshape((N, N + 1), order="F")
    assert np.all(np.abs(lx) <= 0.5 * np.spacing(np.abs(Bx[:, :, 0])))
    assert np.all(np.abs(ly) <= 0.5 * np.spacing(np.abs(By[:, :, 0])))
    assert np.abs(lo).max() > 0          # the CT update has carried low parts
    p = pkg.Solver(32, 32, flux_type=pkg.HLLD, M=50.0, CFL=0.2, pp=1, indicator=pkg.IND_KXRCF, Ck=1.0)
    p.load_checkpoint(str(path))
    uq, qx, qy = p.download()
    assert np.array_equal(qx, Bx) and np.array_equal(qy, By)
    assert p.get_time()[2] == 3


def test_numerical_flags(pkg):
    """the positive statuses: a non-finite state, a negative pressure and a
    collapsed dt are reported (the reference carries on silently), outputs
    are still filled, and mhd2d_set_time clears the flags"""
    s = pkg.Solver(16, 16, flux_type=pkg.HLLD, M=50.0, CFL=0.2, pp=0, indicator=pkg.IND_KXRCF, Ck=1.0)
    s.init_data(pkg.IC_OT_PERTURBED, [1e-2, 3])
    s.rk3_prologue()
    s.step()
    d = s.diagnostics()
    assert d["flags"] == 0 and s.flags == 0
    uh, Bx, By = s.download()
    bad = uh.copy()
    bad[5, 5, 0, 0] = np.nan                       # one cell's density
    s.upload(bad, Bx, By)
    t = s.step(sync=True)
    assert s.flags & pkg.FLAG_NONFINITE and np.isfinite(t)
    s.set_time(0.0)
    assert s.flags == 0 and s.get_time()[2] == 0
    neg = uh.copy()
    neg[7, 7, 0, 4] = 0.1 * neg[7, 7, 0, 4] - 5.0  # energy below the kinetic + magnetic part
    s.upload(neg, Bx, By)
    d = s.diagnostics()
    assert d["pmin"] < 0 and d["flags"] & pkg.FLAG_NEG_PRESSURE
    s.set_time(0.0)
    big = uh.copy()
    big[3, 3, 0, 5] = 1e12                         # a huge Bx: the fast speed collapses the CFL step
    s.upload(big, Bx, By)
    s.step(sync=True)
    assert s.flags & pkg.FLAG_DT_SMALL
