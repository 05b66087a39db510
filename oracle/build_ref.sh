#!/bin/bash
# Recipe for building the reference implementation of the hot path from its
# own sources under /root/reference into oracle/_ref/ (outputs only there).
#
# The path (src/3d/FFT_analysis3D.cpp, src/3dTherm/FFT_thermal3D.cpp) includes
# <eigen/Eigen/Dense>, fftw3.h and VTK headers (include/FFT_analysis3D.hpp:5-23)
# and links FFTW3 + VTK.  None of these exist in this image, so the reference
# is unbuildable here; this script checks for them and stops without writing
# stand-ins.  The oracle (m3d_oracle.c) is pinned without it (tests/test_oracle.py).
set -u
REF=/root/reference
OUT="$(cd "$(dirname "$0")" && pwd)/_ref"
if [ ! -d "$REF" ]; then
    echo "build_ref: $REF absent; nothing to build"
    exit 0
fi
missing=()
echo '#include <eigen/Eigen/Dense>' | g++ -x c++ -fsyntax-only -I"$REF/include" - >/dev/null 2>&1 || missing+=("Eigen (<eigen/Eigen/Dense>)")
echo '#include <fftw3.h>' | g++ -x c++ -fsyntax-only - >/dev/null 2>&1 || missing+=("FFTW3 (fftw3.h)")
echo '#include <vtkVersion.h>' | g++ -x c++ -fsyntax-only - >/dev/null 2>&1 || missing+=("VTK (vtkVersion.h)")
if [ ${#missing[@]} -ne 0 ]; then
    echo "build_ref: reference unbuildable here, missing: ${missing[*]}"
    exit 0
fi
mkdir -p "$OUT"
g++ -O2 -std=c++14 -fPIC -shared -I"$REF/include" -o "$OUT/libmat3d_ref.so" \
    "$REF/src/3d/FFT_analysis3D.cpp" "$REF/src/3dTherm/FFT_thermal3D.cpp" -lfftw3 -lfftw3_threads -lm
