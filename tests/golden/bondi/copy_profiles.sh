#!/bin/sh
# Copies the reference's analytic Bondi profiles (data files, unchanged) into
# this directory: initbounddata/bondigrid<N>.dat for N = 16 ... 512, the
# Gamma = 4/3, r_c = 8 flow tabulated at the zone centres of an N-zone
# log-radius grid from Rin = 1.9 to Rout = 20 (11 ghost zones each side).
set -e
SRC=${1:-/root/reference/initbounddata}
DST=$(dirname "$0")
for n in 16 32 64 128 256 512; do
  cp "$SRC/bondigrid$n.dat" "$DST/bondigrid$n.dat"
done
