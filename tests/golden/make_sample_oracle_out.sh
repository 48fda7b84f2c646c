#!/bin/sh
# Regenerates tests/golden/sample_oracle_out.txt: the oracle's period_search_out for the reference's
# sample input (tests/golden/sample_period_search_in, a copy of period_search/period_search_in).
# Line 1 runs the convexity-weight search; the rest use its result, split over 8 processes.
set -e
cd "$(dirname "$0")/../.."
make -C oracle >/dev/null
IN=tests/golden/sample_period_search_in
./oracle/ps_oracle_cli scan $IN 1 1 > /tmp/sample_1.txt  # runs the convexity-weight search
CW=$(python3 -c "import sys; sys.path.insert(0,'.'); from tests import oracle_binding as ob; print('%.17g' % ob.OracleCtx.from_file('$IN').conw_r0)")
for k in 0 1 2 3 4 5 6 7; do a=$((2 + k*17)); b=$((a+16)); [ $b -gt 134 ] && b=134
  ./oracle/ps_oracle_cli scan $IN $a $b $CW > /tmp/sample_part_$k.txt & done; wait
cat /tmp/sample_1.txt /tmp/sample_part_[0-7].txt > tests/golden/sample_oracle_out.txt
# first 13 lines again with the FMA-contracted build (a second valid rounding)
./oracle/ps_oracle_cli_fma scan $IN 1 1 $CW > /tmp/sample_fma_1.txt
./oracle/ps_oracle_cli_fma scan $IN 2 13 $CW > /tmp/sample_fma_2.txt
cat /tmp/sample_fma_1.txt /tmp/sample_fma_2.txt > tests/golden/sample_oracle_fma_out.txt
