#!/bin/bash
# Builds the reference's own Fortran programs (icosprocess, local_refine)
# from the sources where they lie under /root/reference, plus the unit
# drivers in oracle/ref_drivers/ that call individual reference routines.
# Everything it writes goes to oracle/_ref/ (git-ignored).  Nothing here is
# copied into the repository.  Needs amdflang (ROCm) and MKL's FFTW3
# wrappers (libmkl_rt from /opt/conda/lib).
set -euo pipefail
REF=${CIM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-amdflang}
MKL=${MKL_DIR:-/opt/conda/lib}
[ -d "$REF/icosprocess/SRC" ] || { echo "reference sources not found at $REF"; exit 3; }
mkdir -p "$OUT/icos" "$OUT/lref"
FFLAGS="-O2"
LIBS="-L$MKL -lmkl_rt -Wl,-rpath,$MKL"

build_objs() {   # $1 = object dir, rest = sources (modules first)
  local od=$1; shift
  for f in "$@"; do
    local o="$od/$(basename "${f%.f90}").o"
    if [ ! -f "$o" ] || [ "$f" -nt "$o" ]; then
      $FC $FFLAGS -J "$od" -c "$f" -o "$o"
    fi
  done
}

# icosprocess: modules first, then everything else in SRC and icosprocess/
S=$REF/icosprocess/SRC
ICOS_SRCS=("$S/0common_icosproc.f90" "$S/FFTW.f90")
for f in "$S"/*.f90 "$REF"/icosprocess/icosprocess/*.f90; do
  case "$(basename "$f")" in 0common_icosproc.f90|FFTW.f90) ;; *) ICOS_SRCS+=("$f");; esac
done
build_objs "$OUT/icos" "${ICOS_SRCS[@]}"
ICOS_LIBOBJS=$(ls "$OUT"/icos/*.o | grep -v '/0main.o$')
$FC $FFLAGS -o "$OUT/icosprocess" $ICOS_LIBOBJS "$OUT/icos/0main.o" $LIBS

# icosprocess_fsc: the same objects with a main whose 'FSC' key is parsed.
# The shipped 0main.f90:155 compares the key with 'fsc' while its key table
# holds 'FSC' (0input_parameters.f90:30), so FSC=y never reaches
# reconstruct_half / reconstruct_FSC.  The patched main is a generated file
# under _ref/ only, made to record golden vectors of that otherwise
# unreachable path.
mkdir -p "$OUT/icos_fsc"
sed "s/\.eq\.'fsc')/.eq.'FSC')/" "$S/0main.f90" > "$OUT/icos_fsc/0main_fsckey.f90"
if ! grep -q "eq.'FSC')" "$OUT/icos_fsc/0main_fsckey.f90"; then echo "FSC key patch did not apply"; exit 4; fi
$FC $FFLAGS -J "$OUT/icos" -c "$OUT/icos_fsc/0main_fsckey.f90" -o "$OUT/icos_fsc/0main_fsckey.o"
$FC $FFLAGS -o "$OUT/icosprocess_fsc" $ICOS_LIBOBJS "$OUT/icos_fsc/0main_fsckey.o" $LIBS

# icosprocess_corrctf: the same objects with a main whose 'corrCTF' key is
# parsed.  0main.f90:171 compares the key with 'corrctf' while the key table
# holds 'corrCTF' (0input_parameters.f90:34), so corrCTF=step,cycles never
# reaches reconstruct_corrctf / reconstruct_corr_ctf_substep (:190-197).  As
# above, the patched main is generated under _ref/ only, to record golden
# vectors of that path.
mkdir -p "$OUT/icos_corrctf"
sed "s/\.eq\.'corrctf')/.eq.'corrCTF')/" "$S/0main.f90" > "$OUT/icos_corrctf/0main_corrctfkey.f90"
if ! grep -q "eq.'corrCTF')" "$OUT/icos_corrctf/0main_corrctfkey.f90"; then echo "corrCTF key patch did not apply"; exit 4; fi
$FC $FFLAGS -J "$OUT/icos" -c "$OUT/icos_corrctf/0main_corrctfkey.f90" -o "$OUT/icos_corrctf/0main_corrctfkey.o"
$FC $FFLAGS -o "$OUT/icosprocess_corrctf" $ICOS_LIBOBJS "$OUT/icos_corrctf/0main_corrctfkey.o" $LIBS

# local_reconstruct: flang rejects one line of 0input_parameters.f90 (:32,
# local_sym='c1', a character constant assigned to an undeclared implicit
# integer that nothing reads).  The program is built from its own sources
# with a generated copy of that file lacking the line, under _ref/ only.
LR=$REF/local_reconstruct/SRC
mkdir -p "$OUT/lrec"
sed "/^ *local_sym='c1'/d" "$LR/0input_parameters.f90" > "$OUT/lrec/0input_parameters_nolocalsym.f90"
if grep -q "local_sym='c1'" "$OUT/lrec/0input_parameters_nolocalsym.f90"; then echo "local_sym line not removed"; exit 4; fi
LREC_SRCS=("$LR/0common_icosproc.f90" "$LR/FFTW.f90")
for f in "$LR"/*.f90; do
  case "$(basename "$f")" in 0common_icosproc.f90|FFTW.f90|0input_parameters.f90) ;; *) LREC_SRCS+=("$f");; esac
done
LREC_SRCS+=("$OUT/lrec/0input_parameters_nolocalsym.f90")
build_objs "$OUT/lrec" "${LREC_SRCS[@]}"
$FC $FFLAGS -o "$OUT/local_reconstruct" "$OUT"/lrec/*.o $LIBS

# local_box (sub-particle coordinates / boxes around a vertex, per operator)
B=$REF/local_box/SRC
mkdir -p "$OUT/lbox"
LBOX_SRCS=("$B/0common_icos_loc_box.f90" "$B/FFTW.f90")
for f in "$B"/*.f90; do
  case "$(basename "$f")" in 0common_icos_loc_box.f90|FFTW.f90) ;; *) LBOX_SRCS+=("$f");; esac
done
build_objs "$OUT/lbox" "${LBOX_SRCS[@]}"
$FC $FFLAGS -o "$OUT/local_box" "$OUT"/lbox/*.o $LIBS

# merge_icosprocess_subrec (tnf merge, plain and FSC-weighted)
M=$REF/merge_icosprocess_subrec/SRC
mkdir -p "$OUT/merge"
build_objs "$OUT/merge" "$M/common_merge_icosprocess_subrec.f90" "$M/FFTW.f90" "$M/FFT1d.f90" "$M/help.f90" \
  "$M/icos2f_sym_mat.f90" "$M/img3d_icosavg.f90" "$M/mrchead.f90" "$M/merge_reconstruct.f90" \
  "$M/merge_reconstruct_FSC.f90" "$M/main.f90"
$FC $FFLAGS -o "$OUT/merge_icosprocess_subrec" "$OUT"/merge/*.o $LIBS

# local_refine
L=$REF/local_refine/SRC
LREF_SRCS=("$L/0common_loc_refine.f90" "$L/FFTW.f90")
for f in "$L"/*.f90; do
  case "$(basename "$f")" in 0common_loc_refine.f90|FFTW.f90) ;; *) LREF_SRCS+=("$f");; esac
done
build_objs "$OUT/lref" "${LREF_SRCS[@]}"
LREF_LIBOBJS=$(ls "$OUT"/lref/*.o | grep -v '/0main.o$')
$FC $FFLAGS -o "$OUT/local_refine" $LREF_LIBOBJS "$OUT/lref/0main.o" $LIBS

# unit drivers (own code) linked against the reference objects
for d in "$HERE"/ref_drivers/icos_*.f90; do
  [ -f "$d" ] || continue
  $FC $FFLAGS -J "$OUT/icos" -o "$OUT/$(basename "${d%.f90}")" "$d" $ICOS_LIBOBJS $LIBS
done
for d in "$HERE"/ref_drivers/lref_*.f90; do
  [ -f "$d" ] || continue
  $FC $FFLAGS -J "$OUT/lref" -o "$OUT/$(basename "${d%.f90}")" "$d" $LREF_LIBOBJS $LIBS
done
echo "reference built into $OUT"
