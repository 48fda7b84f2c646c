#!/bin/bash
# Copies a sample of the reference's real period_search_in files (period_search/input_period_search/,
# 100 sparse-survey inputs, Lmax = 3, nrows = 6) into this folder as test fixtures: the inputs with the
# most lightcurves and a few single-curve ones.  Run once where the reference checkout exists.
set -e
SRC=${1:-/root/reference/period_search/input_period_search}
DST=$(dirname "$0")
for f in $(cd "$SRC" && for g in *; do echo "$(sed -n 15p "$g" | awk '{print $1}') $g"; done | sort -rn | head -8 | awk '{print $2}') \
         period_search_in_20188 period_search_in_65646 period_search_in_10537 period_search_in_16686; do
    cp "$SRC/$f" "$DST/$f"
done
ls "$DST" | grep -c period_search_in
