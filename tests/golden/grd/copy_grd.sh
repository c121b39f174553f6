#!/bin/bash
# Copies the mesh files used by tests/test_grd.py and tests/test_gpu_grd.py from the reference's own test decks
# (data files, unchanged).  Run from the repository root where the reference is
# checked out at /root/reference.
set -e
R=/root/reference/Testing/tri_hp
D=$(dirname "$0")
cp "$R/ins/Convergence/Inputs/square1.grd" "$D/square1.grd"
cp "$R/ins/Spatial Accuracy/Inputs/square4.grd" "$D/square4.grd"
cp "$R/ins/Friction_Slip/Inputs/channel.grd" "$D/channel.grd"
cp "$R/conv-diff/Correctness/Inputs/tim.1.grd" "$D/tim.1.grd"
