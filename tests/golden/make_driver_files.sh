#!/bin/bash
# Regenerates tests/golden/driver_files/ with the Fortran compiler of ROCm
# (amdflang): the Fortran runtime's own formatting is the reference.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
tmp="$(mktemp -d)"
trap 'rm -rf "$tmp"' EXIT
"${FC:-/opt/rocm/bin/amdflang}" "$here/driver_files.f90" -o "$tmp/driver_files"
(cd "$tmp" && ./driver_files)
mkdir -p "$here/driver_files"
cp "$tmp"/efmt.txt "$tmp"/go2d_00007_00001.dat "$tmp"/tmask_00001.dat "$tmp"/tmask_raw_00001.dat \
   "$here/driver_files/"
