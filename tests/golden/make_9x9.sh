#!/bin/bash
# Golden vectors for the 9x9-capable symmetry DFS (experiments/9x9/sym_fold9.cpp):
# the reference's binary built from its own source (`make -C oracle ref` ->
# oracle/_ref/sym_fold9), run with MF_STREAM_CASES=1 on a few windows.
# Writes tests/golden/sym9.json.  About 4 minutes on 8 cores.
set -euo pipefail
cd "$(dirname "$0")/../.."
BIN=oracle/_ref/sym_fold9
[ -x "$BIN" ] || make -s -C oracle ref
out=tests/golden/sym9.json
rows=()
win() {   # n expanded max start
    local ans
    ans=$(MF_STREAM_CASES=1 MF_EXPANDED_N=$2 "$BIN" "$1" 8 "$3" "$4" 2>/dev/null)
    rows+=("{\"n\": $1, \"expanded\": $2, \"max\": $3, \"start\": $4, \"answer\": $ans}")
}
win 9 70 50 0
win 9 64 50 0
win 9 60 200 0
win 9 64 1000 5000
win 9 56 20 0
win 9 64 300 123456
win 9 50 1000 0
win 9 48 200 0
win 9 44 50 0
win 9 44 2000 0
n5=$("$BIN" 5 8 2>/dev/null)
n6=$("$BIN" 6 8 2>/dev/null)
setup=$(MF_SETUP_ONLY=1 "$BIN" 9 8 2>&1 >/dev/null | grep -E '^center_size=' | sed 's/center_size=\([0-9]*\) center_cases=\([0-9]*\)/\1 \2/')
reduced=$(MF_SETUP_ONLY=1 "$BIN" 9 8 2>&1 >/dev/null | grep -E '^reduced_cases=' | cut -d= -f2)
{
    echo '{"_source": "oracle/_ref/sym_fold9 (reference experiments/9x9/sym_fold9.cpp built from source), MF_STREAM_CASES=1, threads 8; stream totals at depth 25/27 published in experiments/9x9/README.md",'
    echo " \"squares\": {\"5\": $n5, \"6\": $n6},"
    set -- $setup
    echo " \"center_9\": {\"center_size\": $1, \"center_cases\": $2, \"reduced_cases\": $reduced, \"multiplier_hist\": {\"8\": 1333, \"16\": 464550}},"
    echo ' "stream_totals_9": {"25": 465883, "27": 3157223},'
    echo ' "windows": ['
    for i in "${!rows[@]}"; do
        sep=","; [ "$i" -eq $((${#rows[@]} - 1)) ] && sep=""
        echo "  ${rows[$i]}$sep"
    done
    echo ' ]}'
} > "$out"
echo "wrote $out"
