#!/bin/bash
# TEST INFRASTRUCTURE: builds tests/sanitize/driver (host C++ of the library and the CPU
# oracle) with AddressSanitizer and UndefinedBehaviorSanitizer, host compiler only.
# usage: build.sh <output binary>
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
C=$ROOT/boundaryintegralequations.jl_amd/csrc
O=$ROOT/oracle
OUT=${1:-$HERE/driver}
FLAGS="-std=gnu++17 -O1 -g -fno-omit-frame-pointer -ffp-contract=off -fsanitize=address,undefined -fno-sanitize-recover=undefined"
g++ $FLAGS -fopenmp -I/opt/rocm/include -D__HIP_PLATFORM_AMD__ -I$C \
  $HERE/driver.cpp $C/schedule.cpp $C/rules.cpp $C/near_field.cpp \
  $O/select_exact.cpp $O/bie_oracle.cpp $O/sparse_oracle.cpp -o "$OUT"
