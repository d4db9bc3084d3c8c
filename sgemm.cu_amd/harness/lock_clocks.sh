#!/bin/bash
# Lock the GPU engine clock for repeatable GEMM timings: the rocm-smi
# counterpart of the reference's scripts/lock_clocks.sh:3-6 (which locks the
# base graphics clock and the top memory clock with nvidia-smi).
#
# AMD Instinct parts expose "performance determinism": a fixed ceiling on the
# engine clock (SCLK) that the firmware holds under load, so kernels do not
# ride the boost/thermal curve.  HBM runs at a fixed clock on MI355X, so there
# is no memory clock to lock.  There is no persistence mode to enable: the
# amdgpu driver keeps the device initialised while the system is up, which is
# what the reference's enable_pmode.sh / disable_pmode.sh toggle on NVIDIA.
#
#   sudo ./lock_clocks.sh [SCLK_MHZ] [GPU_INDEX]
#
# SCLK_MHZ defaults to 2100 (below the 2400 MHz peak, which the part cannot
# hold through a long FP32 or BF16 matrix run).  Needs root.  Undo with
# reset_clocks.sh.
set -euo pipefail
sclk="${1:-2100}"
gpu="${2:-0}"
rocm-smi -d "$gpu" --setperfdeterminism "$sclk"
rocm-smi -d "$gpu" --showclocks --showperflevel
