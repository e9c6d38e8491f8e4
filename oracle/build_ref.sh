#!/bin/bash
# build_ref.sh -- compile the reference MCM time-step path from the read-only
# reference sources, in place, into oracle/_ref/ (git-ignored).
#
# TEST INFRASTRUCTURE ONLY: the result is a checker (and the timed CPU
# baseline), never part of the product path.
#
# Recipe:
#   * ROCm flang compiles the reference's own .f90 files listed below (the
#     module database, the material set-up, the initialisation chain and every
#     routine of the SPH step) from /root/reference/MCM, -O2, no -march, so no
#     FMA contraction is possible on x86-64.
#   * They are linked into oracle/_ref/libmcmref.so.  Two symbols stay
#     unresolved and are bound lazily: mcm_shutdown_ (the fatal-error exit;
#     its file needs Intel time/ctime/etime, which flang does not provide) and
#     mcm_transducer_nbr_ (only called when pressure transducers are defined).
#     No stand-in is written for either: reaching them aborts the probe with a
#     loader error.
#   * oracle/ref_probe.f90 (project-written) is compiled against the module
#     files and linked to the library.
#
# Usage: oracle/build_ref.sh   (no-op with a message when /root/reference is absent)
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REF=/root/reference/MCM
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
if [ ! -d "$REF" ]; then
  echo "build_ref: $REF not present; reference probe not built"
  exit 0
fi
if [ ! -x "$FC" ]; then
  echo "build_ref: flang not found at $FC; reference probe not built"
  exit 0
fi
mkdir -p "$OUT/obj" "$OUT/mod"
FILES="mcm_database mcm_math_func
mcm_sets1 mcm_sets3 mcm_sets4 mcm_sets10 mcm_printm
mcm_initialise_database mcm_init_boundcond mcm_calcmass mcm_initvbls mcm_calcg mcm_calcrhopc mcm_eoscalc
mcm_eos1p mcm_eos4p mcm_eos13p mcm_eos28p mcm_eos41p
mcm_eos1c mcm_eos4c mcm_eos13c mcm_eos28c mcm_eos41c
mcm_initsigma mcm_initrod mcm_init_h mcm_aviscosity mcm_nodalav mcm_init_neighb
mcm_init_boundary mcm_calccritts mcm_calctotale mcm_initold
mcm_neighbours mcm_setuplist mcm_ghost_setup mcm_ll_neighbours
mcm_strain mcm_oldstrain mcm_corr_strain mcm_kernel mcm_gradw mcm_get_j_strain_info
mcm_update_h mcm_constitutive mcm_rhoupd mcm_stress_rot1
mcm_f3dm1 mcm_f3dm3 mcm_f3dm9 mcm_f3dm10 mcm_f3dm25
mcm_lieupd mcm_hieupd mcm_sueos mcm_eqos mcm_eqos1 mcm_eqos4 mcm_eqos13 mcm_eqos28 mcm_eqos41
mcm_stress_up mcm_momentum mcm_update_ghost_stress mcm_oldacceleration mcm_corr_acceleration
mcm_get_j_moment_info mcm_boundcond mcm_update_velocity mcm_move mcm_check_sym_pen
mcm_rep_contact mcm_get_j_cont_info
mcm_gttxsg mcm_matin mcm_nodein mcm_nodehist mcm_Inputvel mcm_inp_symmetry_v2 mcm_contactype mcm_baseaccel_in
mcm_termin mcm_transducers_in"
OBJS=""
# mcm_f3dm25.f90 (the cap model) reads locals it never assigns (capa at
# :205-206, blk and davg at :182); it is compiled as the reference's Debug
# configuration compiles every file (MCM.vfproj: /Qsave), where locals are
# static and zero until assigned: flang's -fno-automatic.
for f in $FILES; do
  src="$REF/$f.f90"
  obj="$OUT/obj/$f.o"
  extra=""
  if [ "$f" = "mcm_f3dm25" ]; then extra="-fno-automatic"; fi
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ] || [ "$0" -nt "$obj" ]; then
    "$FC" -O2 -fPIC $extra -module-dir "$OUT/mod" -c "$src" -o "$obj"
  fi
  OBJS="$OBJS $obj"
done
"$FC" -shared -fPIC -Wl,-z,lazy $OBJS -o "$OUT/libmcmref.so"
"$FC" -O2 -fPIC -I "$OUT/mod" -module-dir "$OUT/mod" -c "$HERE/ref_deck.f90" -o "$OUT/obj/ref_deck.o"
"$FC" -O2 -I "$OUT/mod" -module-dir "$OUT/mod" "$HERE/ref_probe.f90" "$OUT/obj/ref_deck.o" -o "$OUT/ref_probe" \
  -L"$OUT" -lmcmref -Wl,--allow-shlib-undefined -Wl,-z,lazy -Wl,-rpath,'$ORIGIN'
echo "build_ref: built $OUT/ref_probe"

# The reference's own time loop and initialisation with its writers
# (libmcmloop.so), driven by dropin_main.f90 twice:
#   ref_solution: with every step routine of the reference (libmcmref.so);
#   mcm_dropin:   with the GPU wrappers of mcmtop_amd/fortran/mcm_gpu_dropin.f90
#                 in place of the step routines they replace (libmcmrest.so
#                 holds the reference's objects minus those), on
#                 mcmtop_amd/libmcm_gpu.so.
LOOP="mcm_solution mcm_initial mcm_init_state_output mcm_state_output mcm_inittimehist mcm_timehist
mcm_geometry mcm_write_ensight_3d mcm_write_res mcm_write_scalar mcm_write_vector mcm_write_tensor"
REPLACED=" mcm_setuplist mcm_init_boundary mcm_ghost_setup mcm_ll_neighbours mcm_strain mcm_update_h
mcm_constitutive mcm_rhoupd mcm_calccritts mcm_rep_contact mcm_momentum mcm_boundcond mcm_calctotale
mcm_update_velocity mcm_move "
LOBJS=""
for f in $LOOP; do
  src="$REF/$f.f90"
  obj="$OUT/obj/$f.o"
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
    "$FC" -O2 -fPIC -I "$OUT/mod" -module-dir "$OUT/mod" -c "$src" -o "$obj"
  fi
  LOBJS="$LOBJS $obj"
done
"$FC" -shared -fPIC -Wl,-z,lazy $LOBJS -o "$OUT/libmcmloop.so"
REPLACED=" $(echo $REPLACED) "
ROBJS=""
for f in $FILES; do
  case "$REPLACED" in *" $f "*) continue ;; esac
  ROBJS="$ROBJS $OUT/obj/$f.o"
done
"$FC" -shared -fPIC -Wl,-z,lazy $ROBJS -o "$OUT/libmcmrest.so"
"$FC" -O2 -I "$OUT/mod" -module-dir "$OUT/mod" "$HERE/dropin_main.f90" "$OUT/obj/ref_deck.o" -o "$OUT/ref_solution" \
  -L"$OUT" -lmcmloop -lmcmref -Wl,--allow-shlib-undefined -Wl,-z,lazy -Wl,-rpath,'$ORIGIN'
echo "build_ref: built $OUT/ref_solution"
GPULIB="$HERE/../mcmtop_amd"
FDIR="$HERE/../mcmtop_amd/fortran"
if [ -f "$GPULIB/libmcm_gpu.so" ]; then
  "$FC" -O2 -fPIC -I "$OUT/mod" -module-dir "$OUT/mod" -c "$FDIR/mcm_gpu_iface.f90" -o "$OUT/obj/mcm_gpu_iface.o"
  "$FC" -O2 -fPIC -I "$OUT/mod" -module-dir "$OUT/mod" -c "$FDIR/mcm_gpu_dropin.f90" -o "$OUT/obj/mcm_gpu_dropin.o"
  "$FC" -O2 -I "$OUT/mod" -module-dir "$OUT/mod" -rdynamic "$HERE/dropin_main.f90" "$OUT/obj/ref_deck.o" \
    "$OUT/obj/mcm_gpu_iface.o" "$OUT/obj/mcm_gpu_dropin.o" -o "$OUT/mcm_dropin" \
    -L"$OUT" -lmcmloop -lmcmrest -L"$GPULIB" -lmcm_gpu -Wl,--allow-shlib-undefined -Wl,-z,lazy \
    -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/../../mcmtop_amd'
  echo "build_ref: built $OUT/mcm_dropin"
fi
