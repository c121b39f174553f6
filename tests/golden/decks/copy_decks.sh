#!/bin/bash
# Copies the input decks used by tests/test_input_map.py from the reference's own test
# decks (data files, unchanged).  Run from the repository root where the reference is
# checked out at /root/reference.
set -e
R=/root/reference/Testing/tri_hp
D=$(dirname "$0")
cp "$R/ins/Convergence/Inputs/run_master.inpt" "$D/ins_convergence_run_master.inpt"
cp "$R/ins/Restart Test/Inputs/run.inpt" "$D/ins_restart_test_run.inpt"
