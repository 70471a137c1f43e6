#!/bin/sh
# Writes the golden per-graph results of config 1 (sincqed run 1.0 97 1e-12 all2) from the reference
# implementation itself (oracle/_ref/sincqed_ref, built from the reference sources by oracle/Makefile):
# config1_default.txt (symmetry fold on, the reference's default) and config1_sym0.txt (fold off, the
# semantics of the GPU path), each the reference's results lines plus its full-precision "#R" records.
set -e
cd "$(dirname "$0")"
R=../../oracle/_ref/sincqed_ref
for mode in default sym0; do
  out=config1_$mode.txt; rm -f $out.res
  if [ $mode = sym0 ]; then export SINCQED_SYM=0; else unset SINCQED_SYM; fi
  SINCQED_CKPT=0 SINCQED_PROGRESS=0 SINCQED_RESULTS=$out.res OMP_NUM_THREADS=4 $R run 1.0 97 1e-12 all2 | grep '^#R' > $out.R
  grep -v '^# [0-9][0-9][0-9][0-9]-' $out.res | grep -v '^# window' > $out; cat $out.R >> $out; rm -f $out.res $out.R
done
