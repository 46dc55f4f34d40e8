#!/bin/bash
# Golden vectors for sym_fold's materialised mode with MF_COST_DEPTH /
# MF_SORT_CASES (sym_fold.cpp:267-314): the reference's own binary, built from
# its sources by `make -C oracle ref` (oracle/_ref/sym_fold), run on a few
# scored and sorted windows.  Writes tests/golden/sym_sorted_windows.json.
# Takes about 3 minutes on 8 cores (the 7x7 windows hold large cases).
set -euo pipefail
cd "$(dirname "$0")/../.."
BIN=oracle/_ref/sym_fold
[ -x "$BIN" ] || make -s -C oracle ref
out=tests/golden/sym_sorted_windows.json
rows=()
run() {   # n max start depth sort
    local n=$1 max=$2 start=$3 depth=$4 sort=$5 err ans
    err=$(mktemp)
    ans=$(MF_COST_DEPTH=$depth MF_SORT_CASES=$sort "$BIN" "$n" 8 "$max" "$start" 2>"$err")
    local cost
    cost=$(grep '^cost_depth=' "$err" | sed 's/ score_s=.*//')
    rows+=("{\"n\": $n, \"max\": $max, \"start\": $start, \"cost_depth\": $depth, \"sort\": \"$sort\", \"answer\": $ans, \"cost_line\": \"$cost\"}")
    rm -f "$err"
}
run 7 2000 1000 3 1
run 7 1500 0 3 asc
run 7 500 7 2 1
run 6 300 100 4 1
run 6 0 0 3 ""
{
    echo '{"_source": "oracle/_ref/sym_fold (reference sym_fold.cpp built from source), MF_COST_DEPTH/MF_SORT_CASES, threads 8",'
    echo ' "items": ['
    for i in "${!rows[@]}"; do
        sep=","; [ "$i" -eq $((${#rows[@]} - 1)) ] && sep=""
        echo "  ${rows[$i]}$sep"
    done
    echo ' ]}'
} > "$out"
echo "wrote $out"
