#!/usr/bin/env bash
# build_ref.sh -- compile the reference Fortran (GDF_Balsara_P2_0316) with
# flang into oracle/_ref/, for use as the checker of the CPU restatement and
# as the "reference" CPU baseline.  Runs only where /root/reference exists.
#
# The reference is never copied into this repository: each variant is built
# in a throw-away directory under /tmp and only the linked binary lands in
# oracle/_ref/ (git-ignored).  The reference fixes its configuration at
# compile time, so each variant applies these configuration edits to the
# scratch copy (the same edits the author makes by hand):
#   P1 com.txt:1            Nx, Ny, flux_type; flash = 0 (no frame files)
#   P2 init_data.f90:19     one object per initial condition (init1 vortex,
#                           init3 Orszag-Tang, init4 rotor, init5 blast,
#                           init6/6b cloud-shock, init7 magnetic loop,
#                           init8/8b Brio-Wu, init9/9b Ryu-Jones), subroutine
#                           renamed init_data_<N>.  init0 and init2 define six
#                           conserved variables (an older equation set) and do
#                           not compile with this init_data
#   P5 TVB_Limiter_P2.f90:123  'call pp_Limiter' enabled (variants *_pp)
#   P7 HLL_Flux_2D.f90:8    EzRU..EzLD are read uninitialised at :61-62;
#                           they are set to the corner fields v*B1 - u*B2
#                           (calculate_Ez, as LF_Flux_2D.f90:18-21 does)
#   P6 KXRCF_Detector.f90   does not compile as shipped (uG undeclared at
#                           :17,91): uG is declared as a saved local of the
#                           shape its loops index; Ck (:7, shipped 0, which
#                           flags every cell) is set to the variant's value;
#                           Is_Trouble_Cell is reset at entry (the shipped
#                           code never clears a flag).  In variants *_kx the
#                           commented KXRCF gating of TVB_Limiter_P2.f90:12,
#                           22,116 is enabled.
# main.f90 is replaced by oracle/ref_driver.f90 in these variants.
#
# The drop-in pair (build_main below) keeps the reference's own main.f90 and
# RK3.f90 untouched, with frame output on (com.txt:1 flash = 1, frameMAX
# set), the Orszag-Tang include (init_data.f90:19 init3.txt) and a shorter
# tend (init3.txt:49): ref_main_<name> links the reference's hot-path
# sources, dropin_<name> links freljord-cfd-arsenal_amd/fortran/
# mhd2d_com_shims.f90 in their place, over libmhd2d.so (its mhd2d_capi
# module comes from the package's own build, make -C
# freljord-cfd-arsenal_amd fortran).
set -euo pipefail

REF=${REF:-/root/reference/GDF_Balsara_P2_0316}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-flang}
JOBS=${JOBS:-8}

if [ ! -d "$REF" ]; then
  echo "build_ref.sh: $REF not present; reference build skipped" >&2
  exit 0
fi
mkdir -p "$OUT"

# name Nx Ny flux_type pp openmp kxrcf-gating (Ck = 1 everywhere)
VARIANTS=${VARIANTS:-"
r16_f4 16 16 4 0 0 0
r16_f2 16 16 2 0 0 0
r16_f1 16 16 1 0 0 0
r16_f3 16 16 3 0 0 0
r16_f4_pp 16 16 4 1 0 0
r12x8_f4 12 8 4 0 0 0
r32_f4_pp 32 32 4 1 0 0
r32_f2 32 32 2 0 0 0
r32_f4_pp_kx 32 32 4 1 0 1
r32_f2_kx 32 32 2 0 0 1
r512_f4_pp_kx_omp 512 512 4 1 1 1
"}
CK=${CK:-1}

SRCS="get_GLP get_basis set_bc Lh LhEz eigenvalueMm HLL_Flux LF_Flux HLLC_Flux HLLD_Flux
HLL_Flux_2D LF_Flux_2D HLLC_Flux_2D HLLD_Flux_2D HLL1 MHD_flux calculate_Ez TVB_Limiter_P2
minmod minmodB eigenmatrix pp_Limiter pressure div_free_Balsara calculate_dt calculate_umax
calculate_pmin calculate_totaldiv RK3 Rotor_functions KXRCF_Detector p0 Cloud_Shock_functions
Cloud_Shock_functions2 Loop_functions BrioWu_functions RyuJones_functions calculate_Az calculate_L2error"
ICS="3 1 4 5 6 6b 7 8 8b 9 9b"

build_one() {
  local name=$1 nx=$2 ny=$3 ft=$4 pp=$5 omp=$6 kx=$7
  local w
  w=$(mktemp -d /tmp/gdf_ref_XXXXXX)
  cp "$REF"/*.f90 "$REF"/*.txt "$w"/
  chmod -R u+w "$w"
  # P1
  sed -i "1s/Nx = [0-9]*, Ny = [0-9]*/Nx = $nx, Ny = $ny/; 1s/flux_type = [0-9]*/flux_type = $ft/; 1s/flash = [0-9]*/flash = 0/" "$w/com.txt"
  # P2
  for ic in $ICS; do
    sed -e "s/include 'init6.txt'/include 'init$ic.txt'/" \
        -e "s/subroutine init_data\b/subroutine init_data_$ic/" \
        "$w/init_data.f90" > "$w/init_data_$ic.f90"
  done
  # P5
  if [ "$pp" = 1 ]; then
    sed -i "s/^\(\s*\)!call pp_Limiter/\1call pp_Limiter/" "$w/TVB_Limiter_P2.f90"
  fi
  # P6
  sed -i -e "/^\s*real Iij(NumEq),Lij,uinfty(NumEq),Ck/a\\
    real, save :: uG(Nx,Ny,NumGLP,NumGLP,NumEq)" \
         -e "s/^\(\s*\)Ck = 0\s*$/\1Ck = $CK\\
\1Is_Trouble_Cell = 0/" "$w/KXRCF_Detector.f90"
  if [ "$kx" = 1 ]; then
    sed -i -e "s/^\(\s*\)!call KXRCF_Detector/\1call KXRCF_Detector/" \
           -e "s/^\(\s*\)!if (Is_Trouble_Cell(i,j) == 1) then/\1if (Is_Trouble_Cell(i,j) == 1) then/" \
           -e "s/^\(\s*\)!end if\s*$/\1end if/" "$w/TVB_Limiter_P2.f90"
  fi
  # P7
  sed -i "/^\s*ByLD = ULD1(7)/a\\
    call calculate_Ez(EzRU,URU1(2)/URU1(1),URU1(3)/URU1(1),URU1(6),URU1(7))\\
    call calculate_Ez(EzLU,ULU1(2)/ULU1(1),ULU1(3)/ULU1(1),ULU1(6),ULU1(7))\\
    call calculate_Ez(EzRD,URD1(2)/URD1(1),URD1(3)/URD1(1),URD1(6),URD1(7))\\
    call calculate_Ez(EzLD,ULD1(2)/ULD1(1),ULD1(3)/ULD1(1),ULD1(6),ULD1(7))" "$w/HLL_Flux_2D.f90"
  cp "$HERE/ref_driver.f90" "$w/"
  local flags="-O2 -fdefault-real-8 -fdefault-double-8"
  local ldflags=""
  if [ "$omp" = 1 ]; then flags="-O3 -fdefault-real-8 -fdefault-double-8 -fopenmp"; ldflags="-fopenmp"; fi
  if [ $((nx * ny)) -gt 250000 ]; then flags="$flags -mcmodel=medium"; ldflags="$ldflags -mcmodel=medium -fuse-ld=bfd"; fi
  (
    cd "$w"
    units="$SRCS ref_driver"
    for ic in $ICS; do units="$units init_data_$ic"; done
    objs=""
    for f in $units; do objs="$objs $f.o"; done
    for f in $units; do echo "$FC $flags -c $f.f90"; done | xargs -P "$JOBS" -I{} sh -c "{}"
    $FC $ldflags $objs -o "$OUT/$name"
  )
  rm -rf "$w"
  echo "built $OUT/$name"
}

echo "$VARIANTS" | while read -r name nx ny ft pp omp kx; do
  [ -z "$name" ] && continue
  build_one "$name" "$nx" "$ny" "$ft" "$pp" "$omp" "${kx:-0}"
done

PKG=$(cd "$HERE/../freljord-cfd-arsenal_amd" && pwd)
build_main() {
  local name=$1 nx=$2 ny=$3 ft=$4 tend=$5 fmax=$6
  local w
  w=$(mktemp -d /tmp/gdf_main_XXXXXX)
  cp "$REF"/*.f90 "$REF"/*.txt "$w"/
  chmod -R u+w "$w"
  sed -i "1s/Nx = [0-9]*, Ny = [0-9]*/Nx = $nx, Ny = $ny/; 1s/flux_type = [0-9]*/flux_type = $ft/; 1s/flash = [0-9]*/flash = 1/; 1s/frameMAX = [0-9]*/frameMAX = $fmax/" "$w/com.txt"
  sed -i "s/include 'init6.txt'/include 'init3.txt'/" "$w/init_data.f90"
  sed -i "s/^\(\s*\)tend = 3\s*$/\1tend = $tend/" "$w/init3.txt"
  grep -q "tend = $tend" "$w/init3.txt"
  local flags="-O2 -fdefault-real-8 -fdefault-double-8"
  (
    cd "$w"
    local host="main RK3 init_data get_GLP get_basis save_XY save_solution calculate_L2error"
    local hot="set_bc Lh LhEz eigenvalueMm HLL_Flux LF_Flux HLLC_Flux HLLD_Flux HLL_Flux_2D LF_Flux_2D HLLC_Flux_2D
      HLLD_Flux_2D HLL1 MHD_flux calculate_Ez TVB_Limiter_P2 minmod minmodB eigenmatrix pp_Limiter pressure
      div_free_Balsara calculate_dt calculate_umax calculate_pmin calculate_totaldiv KXRCF_Detector"
    sed -i -e "/^\s*real Iij(NumEq),Lij,uinfty(NumEq),Ck/a\\
    real, save :: uG(Nx,Ny,NumGLP,NumGLP,NumEq)" KXRCF_Detector.f90
    for f in $host $hot; do echo "$FC $flags -c $f.f90"; done | xargs -P "$JOBS" -I{} sh -c "{}"
    objs=""
    for f in $host $hot; do objs="$objs $f.o"; done
    $FC $objs -o "$OUT/ref_main_$name"
    cp "$PKG/fortran/mhd2d_com_shims.f90" .
    $FC $flags -I "$PKG/_build" -c mhd2d_com_shims.f90
    objs=""
    for f in $host; do objs="$objs $f.o"; done
    $FC $objs mhd2d_com_shims.o "$PKG/_build/mhd2d_capi.o" -L"$PKG" -lmhd2d \
      -Wl,-rpath,'$ORIGIN/../../freljord-cfd-arsenal_amd' -o "$OUT/dropin_$name"
  )
  rm -rf "$w"
  echo "built $OUT/ref_main_$name $OUT/dropin_$name"
}
if [ -f "$PKG/_build/mhd2d_capi.mod" ] && [ -f "$PKG/libmhd2d.so" ]; then
  build_main r16_f4 16 16 4 0.5 10
else
  echo "build_ref.sh: the package's Fortran module is not built; drop-in pair skipped" >&2
fi
