#!/usr/bin/env python3
"""Driver with the reference's run_fem.py interface (run_fem.py:8-96):

    python run_fem.py <run_config.json> [fem_main options]

For every task of the config it checks the prepared mesh under
<working_path>/prepared_meshes/<DIM>D/<task>, creates
<working_path>/fem_results/<DIM>D/<task>/output, and runs the HIP solver
executable (cudafemproject_amd/fem_main) on it, which writes results.vtk
there.  Unlike the reference, the executable really receives the config and
task (the reference's main ignores its argv, fem/src/main.cpp:18-22).
Extra arguments are passed to fem_main (e.g. --precision mixed).
The raw LS-DYNA .k -> prepared-mesh converter (mesh_parser_lib) is not part
of this build: a task without a prepared mesh is reported and skipped.
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
FEM_MAIN = os.path.join(ROOT, "cudafemproject_amd", "fem_main")


def main(argv):
    if len(argv) < 2:
        print("usage: run_fem.py <run_config.json> [fem_main options]")
        return 2
    config_path = argv[1]
    try:
        with open(config_path) as f:
            cfg = json.load(f)
        work_dir = str(cfg["working_path"])
        task_list = cfg["task"]
        dim = str(cfg["DIM"])
    except Exception as e:  # the reference prints and exits the same way
        print("Something wrong with config file: ", e)
        return 0
    if not os.path.exists(FEM_MAIN):
        print(f"{FEM_MAIN} is not built (python -c 'import __graft_entry__ as g; g.build()')")
        return 1
    status = 0
    for task in task_list:
        prep = os.path.join(work_dir, "prepared_meshes", dim + "D", task)
        res = os.path.join(work_dir, "fem_results", dim + "D", task)
        if not (os.path.isdir(prep) and os.listdir(prep)):
            print(f"prepared mesh for {task} not found in {prep}; "
                  "the .k mesh parser is not part of this build")
            status = 1
            continue
        os.makedirs(os.path.join(res, "output"), exist_ok=True)
        print("\nRUN calculation for " + task, flush=True)
        r = subprocess.run([FEM_MAIN, config_path, task, *argv[2:]])
        if r.returncode != 0:
            print(f"calculation for {task} failed ({r.returncode})")
            status = r.returncode
            continue
        print("FINISHED calculation for " + task)
    return status


if __name__ == "__main__":
    sys.exit(main(sys.argv))
