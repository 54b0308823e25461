#!/usr/bin/env python3
"""Driver with the reference's run_fem.py interface (run_fem.py:8-96):

    python run_fem.py <run_config.json> [fem_main options]

For every task of the config: if <working_path>/prepared_meshes/<DIM>D/<task>
is missing or empty, the raw mesh <CudaFEMproject_path>/raw_meshes/<DIM>D/<task>.k
is converted into it (cudafemproject_amd.kparser, the reference's
mesh_parser_lib); then <working_path>/fem_results/<DIM>D/<task>/output is
created and the HIP solver executable (cudafemproject_amd/fem_main) runs on
the task and writes results.vtk.  Unlike the reference, the executable really
receives the config and task (the reference's main ignores its argv,
fem/src/main.cpp:18-22).  Extra arguments are passed to fem_main (e.g.
--precision mixed).  The reference's plotting post-processing and test runner
(postprocessing/, run_tests.py) are not part of this build.
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
FEM_MAIN = os.path.join(ROOT, "cudafemproject_amd", "fem_main")
sys.path.insert(0, ROOT)


def main(argv):
    if len(argv) < 2:
        print("usage: run_fem.py <run_config.json> [fem_main options]")
        return 2
    config_path = argv[1]
    try:
        with open(config_path) as f:
            cfg = json.load(f)
        proj_dir = str(cfg.get("CudaFEMproject_path", ROOT))
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
            raw = os.path.join(proj_dir, "raw_meshes", dim + "D", task + ".k")
            if not os.path.exists(raw):
                print("raw mesh " + task + ".k doesn't exist")
                return 0
            print("RUN parser for " + task, flush=True)
            from cudafemproject_amd.kparser import run_parser
            run_parser(proj_dir, prep, raw, task, dim)
            print("FINISHED parser for " + task)
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
