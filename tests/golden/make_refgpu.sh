#!/bin/sh
# Writes tests/golden/fixtures/refgpu_h<h>_N<N>_eps<eps>.txt: the "#R" result records of the reference driver
# built as its GPU build (oracle/Makefile's `gpu` target: driver/{sum,eval,lattice,checkpoint,io,graphs}.c
# with -DSQ_GPU, FP contraction off, over this library's gpu.h back end src/sfx_gpu.c), for the g3_lbl jobs of
# the benchmark workloads and the serving sweep.  Needs an MI355X and the built library.
#
# g3_lbl is the one integral whose reference CPU build evaluates a differently rounded double half of the
# kernel (Gauss-Legendre nodes as SIMD lanes) than the QREAL half the GPU path evaluates in FP64; these
# records hold the counts of the reference's own GPU path for it.
set -e
cd "$(dirname "$0")/../.."
export SINCQED_SYM=0 SINCQED_CKPT=0 SINCQED_PROGRESS=0 SINCQED_RESULTS=- OMP_NUM_THREADS=${OMP_NUM_THREADS:-16}
for a in "1.0 1 97 1e-12 1" "0.8 0.8 121 1e-12 1" "0.7 0.7 139 1e-12 2" "1.0 1 97 1e-10 1" "0.9 0.9 109 1e-10 1"; do
    set -- $a
    out=tests/golden/fixtures/refgpu_h$2_N$3_eps$4.txt
    echo "# sincqed_ref_gpu run $1 $3 $4 $5 GRAPH  (SINCQED_SYM=0 SINCQED_CKPT=0; the reference driver built with -DSQ_GPU over this library's gpu.h back end, src/sfx_gpu.c; OMP_NUM_THREADS=$OMP_NUM_THREADS)" > $out.tmp
    timeout -k 10 600 oracle/_ref/sincqed_ref_gpu run $1 $3 $4 $5 g3_lbl | grep '^#R' >> $out.tmp
    mv $out.tmp $out
    echo "$out"
done
