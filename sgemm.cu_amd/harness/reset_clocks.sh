#!/bin/bash
# Undo lock_clocks.sh: the rocm-smi counterpart of the reference's
# scripts/reset_clocks.sh:3-4 (nvidia-smi -rgc / -rmc).  Needs root.
#
#   sudo ./reset_clocks.sh [GPU_INDEX]
set -euo pipefail
gpu="${1:-0}"
rocm-smi -d "$gpu" --resetperfdeterminism
rocm-smi -d "$gpu" --resetclocks
rocm-smi -d "$gpu" --showclocks --showperflevel
