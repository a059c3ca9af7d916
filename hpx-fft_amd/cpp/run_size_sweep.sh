#!/bin/bash
# Problem-size sweep of the MI355X shared_loop_2d_gpu harness, with the
# argument order of HPX-FFT's shared-memory size sweep
# (benchmark/sbatch_scripts/run_hpxfft_size_shared.sh:10-15):
#   $1 executable (e.g. hpx-fft_amd/bin/shared_loop_2d_gpu)
#   $2 planning flag (estimate / measure / patient / exhaustive)
#   $3 core count: kept for the argument order; one process drives one GPU,
#      so it is only echoed (the runtime file's n_threads column reads "gpu")
#   $4 first size power, $5 last size power (square grids 2^p x 2^p)
#   $6 runs per size
# Rows are appended to runtimes/runtimes_hpx_shared_loop.txt under the
# current directory; the first run writes the CSV header (--header=true).
# The script stops at the first run that fails and exits with its status.
set -u
if [ $# -ne 6 ]; then
    echo "usage: $0 EXECUTABLE PLAN CORES POW_START POW_STOP RUNS" >&2
    exit 2
fi
EXE=$1
PLAN=$2
CORES=$3
POW_START=$4
POW_STOP=$5
RUNS=$6
echo "size sweep: $EXE plan=$PLAN cores=$CORES (one GPU) 2^$POW_START..2^$POW_STOP x $RUNS runs"
header=true
for (( p=POW_START; p<=POW_STOP; p++ )); do
    n=$((1 << p))
    for (( r=0; r<RUNS; r++ )); do
        "$EXE" --nx=$n --ny=$n --plan="$PLAN" --header=$header > /dev/null
        rc=$?
        if [ $rc -ne 0 ]; then
            echo "run failed: nx=ny=$n plan=$PLAN (status $rc)" >&2
            exit $rc
        fi
        header=false
    done
    echo "done 2^$p x 2^$p ($RUNS runs)"
done
