#!/bin/bash
# Copies the reference's real period_search_in files (period_search/input_period_search/: 100 sparse-survey
# inputs, Lmax = 3, nrows = 6) into this folder as test fixtures.  Run once where the reference checkout
# exists.
set -e
SRC=${1:-/root/reference/period_search/input_period_search}
DST=$(dirname "$0")
cp "$SRC"/period_search_in_* "$DST"/
ls "$DST" | grep -c period_search_in
