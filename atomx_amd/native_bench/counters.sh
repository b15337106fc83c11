#!/bin/bash
# Hardware-counter runs of the benchmark tensor through the native driver
# (no Python in the profiled process).  Per precision mode: a kernel trace,
# then one pass each of FETCH_SIZE, WRITE_SIZE, the MFMA counters, the LDS
# counters and the wave-cycle (stall) counters, each a profiler run of its
# own; every run ends with a 256 MiB device copy that calibrates the byte
# counters.
#
#   python -m atomx_amd.native_bench.build        (once, on the build host)
#   bash atomx_amd/native_bench/counters.sh OUTDIR
#
# Summarise with: python -m atomx_amd.native_bench.summarize OUTDIR
set -o pipefail
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
OUT=$(mkdir -p "$1" && cd "$1" && pwd)
BIN=$ROOT/atomx_amd/_lib/rk_tensor_bench
[ -x "$BIN" ] || { echo "missing $BIN: run python -m atomx_amd.native_bench.build first"; exit 1; }
(cd "$ROOT" && python3 -m atomx_amd.native_bench.write_workload "$OUT/w.bin") || exit 1
cd /tmp && export TMPDIR=/tmp
pass() {  # name, precision, profiler options...
    local name=$1 prec=$2
    shift 2
    echo "== $name (start)"
    # order -2: every (order, K split) the schedule tuner times, 3 + 10 runs each
    timeout -s KILL 120 rocprofv3 "$@" --output-format csv -d "$OUT/$name" -o nb -- \
        "$BIN" "$OUT/w.bin" "$prec" -2 0 10 256 > "$OUT/$name.log" 2>&1
    local rc=$?
    grep -E "^workload|^copy" "$OUT/$name.log"
    if [ $rc -eq 0 ]; then return 0; fi
    # Any other status fails the pass.  The one exception is a SIGSEGV (139)
    # raised after the driver printed its closing line: the program had
    # finished and its counter files were written; the fault is the
    # profiler's own teardown.  A crash before that line is a failure.
    if [ $rc -eq 139 ] && [ "$(grep -v '^$' "$OUT/$name.log" | tail -1)" = "rk_tensor_bench: done" ]; then
        echo "$name: rocprofv3 exited 139 after the driver finished (results kept)"
        return 0
    fi
    echo "$name: rc=$rc"
    return $rc
}
MFMA="SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_VALU_MFMA_MOPS_I8 SQ_VALU_MFMA_BUSY_CYCLES SQ_BUSY_CYCLES GRBM_GUI_ACTIVE"
# LDS activity and conflicts; where the waves' cycles go (8 SQ counters per pass)
LDS="SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_WAIT_INST_LDS SQ_ACTIVE_INST_LDS SQ_LDS_DATA_FIFO_FULL SQ_LDS_ADDR_CONFLICT SQ_WAVE_CYCLES"
STALL="SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_VMEM SQ_ACTIVE_INST_MISC SQ_INSTS_VALU SQ_WAVE_CYCLES"
for prec in fp64 split; do
    pass "$prec-trace" $prec --kernel-trace --stats &&
    pass "$prec-fetch" $prec --pmc FETCH_SIZE &&
    pass "$prec-write" $prec --pmc WRITE_SIZE &&
    pass "$prec-mfma" $prec --pmc $MFMA &&
    pass "$prec-lds" $prec --pmc $LDS &&
    pass "$prec-stall" $prec --pmc $STALL || exit 1
done
