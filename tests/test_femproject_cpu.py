"""CPU tests of the FEMproject-compatible host layer (fem_main / run_fem.py):
run-config keys and defaults (datakeeper.cpp:6-48), prepared-mesh paths,
both stress.txt layouts."""
import json
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT



@pytest.fixture(autouse=True)
def _driver(fem_main):
    global FEM_MAIN
    FEM_MAIN = fem_main


FEM_MAIN = None


def workspace(tmp_path, fixture="test_rect_pcg", dim=2, **cfg):
    prep = tmp_path / "prepared_meshes" / f"{dim}D" / fixture
    prep.mkdir(parents=True)
    for f in ("nodes.txt", "elements.txt", "constraints.txt", "loads.txt", "stress.txt"):
        shutil.copy(os.path.join(GOLDEN, fixture, f), prep / f)
    c = {"working_path": str(tmp_path) + "/", "task": [fixture], "DIM": dim}
    c.update(cfg)
    p = tmp_path / "cfg.json"
    p.write_text(json.dumps(c))
    return str(p)


def dry(cfg, task="test_rect_pcg"):
    r = subprocess.run([FEM_MAIN, cfg, task, "--dry-run"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_static_defaults(tmp_path):
    out = dry(workspace(tmp_path))
    assert "Nodes count = 6" in out and "Elements count = 4" in out
    assert "Boundary edges with b.c. count = 2" in out
    assert "task_type static" in out and "E 2e+07 nu 0.3" in out     # the reader's defaults
    assert out.strip().endswith("fem_results/2D/test_rect_pcg/results.vtk")


def test_dynamic_keys_only_for_dynamic(tmp_path):
    cfg = workspace(tmp_path, youngModulus=3e7, poissonRatio=0.2, task_type="dynamics",
                    rho=1000, dt=2e-4)
    out = dry(cfg)
    # "dynamics" (as several shipped configs say) is not "dynamic" for the reader
    assert "task_type dynamics" in out and "rho 2400 dt 0.001" in out
    cfg = workspace(tmp_path / "b", task_type="dynamic", rho=1000, dt=2e-4, beta2=0.6,
                    wavelet={"WaveletType": "ricker", "freq": 50})
    out = dry(cfg)
    assert "task_type dynamic" in out and "rho 1000 dt 0.0002" in out and "beta2 0.6" in out
    assert "wavelet ricker" in out


def test_bad_config_is_an_error(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"DIM": 2, "task": [')
    r = subprocess.run([FEM_MAIN, str(p), "x", "--dry-run"], capture_output=True, text=True)
    assert r.returncode == 1 and "JSON parse error" in r.stderr


def test_missing_mesh_is_an_error(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"DIM": 3, "working_path": str(tmp_path)}))
    r = subprocess.run([FEM_MAIN, str(p), "nothere", "--dry-run"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_run_fem_reports_missing_raw_mesh(tmp_path):
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"DIM": 2, "working_path": str(tmp_path) + "/", "task": ["nope"],
                             "CudaFEMproject_path": str(tmp_path) + "/"}))
    r = subprocess.run(["python", os.path.join(ROOT, "run_fem.py"), str(p)], capture_output=True,
                       text=True)
    assert r.returncode == 0 and "raw mesh nope.k doesn't exist" in r.stdout   # as run_fem.py:46-48


def test_run_fem_parses_raw_mesh(tmp_path):
    """No prepared mesh: run_fem.py converts raw_meshes/2D/9task_3.k and hands
    the task to the solver (dry run here: no GPU needed)."""
    raw = tmp_path / "proj" / "raw_meshes" / "2D"
    raw.mkdir(parents=True)
    shutil.copy(os.path.join(GOLDEN, "9task_3", "9task_3.k"), raw / "9task_3.k")
    p = tmp_path / "c.json"
    p.write_text(json.dumps({"DIM": 2, "working_path": str(tmp_path / "work") + "/",
                             "CudaFEMproject_path": str(tmp_path / "proj") + "/",
                             "task": ["9task_3"], "youngModulus": 2e7, "poissonRatio": 0.25}))
    r = subprocess.run(["python", os.path.join(ROOT, "run_fem.py"), str(p), "--dry-run"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FINISHED parser for 9task_3" in r.stdout and "Nodes count = 35" in r.stdout
    prep = tmp_path / "work" / "prepared_meshes" / "2D" / "9task_3"
    for f in ("nodes.txt", "elements.txt", "constraints.txt", "loads.txt", "stress.txt"):
        assert (prep / f).exists()


REFERENCE_SHAPED_MAIN = r'''
// The call sequence of the reference's fem/src/main.cpp:10-36 with the
// device set to HIP, against this build's headers (names at global scope).
#include <femfunc.h>
#include <datakeeper.h>
#include <memory>
int main(int argc, char *argv[]) {
  DEVICE_NAME deviceType = HIP;
  dataKeeper dk(argv[1], argv[2]);
  if (dk.getTaskType() == "static")
    gpuCalculateFEM_EbE_vec(deviceType, dk, true);
  else
    gpuCalculateFEM_DYN(deviceType, dk, true);
  ResultsDataKeeper RESdata2(false, false, false, dk.get_nodesCount());
  WriteResults(dk, RESdata2);
  // the PCG entry point on its own (femfunc.cpp:5-20 call sequence)
  std::unique_ptr<ElementsData> ed(ElementsData::setElementsData(deviceType, dk));
  ed->genMask();
  ed->calculateKlocals();
  ed->initFlocals(0.f, dk.getWaveletParams());
  Matrix *u = dk.get_displacements();
  gpuPCG_EbE_vec(*u, deviceType, dk, *ed, true, 1e-8f, true);
  return u->get_numRows() == size_t(dk.get_nodesCount()) ? 0 : 1;
}
'''


def test_reference_shaped_main_compiles_and_links(tmp_path, fem_main):
    """A caller written against the reference's femfunc.h / datakeeper.h
    (global names, reference signatures of gpuPCG_EbE_vec and WriteResults)
    compiles and links against this build with only the device changed."""
    src = tmp_path / "ref_main.cpp"
    src.write_text(REFERENCE_SHAPED_MAIN)
    pkg = os.path.join(ROOT, "cudafemproject_amd")
    fp = os.path.join(pkg, "femproject")
    srcs = [os.path.join(fp, f) for f in sorted(os.listdir(fp))
            if f.endswith(".cpp") and f != "main.cpp"]
    exe = tmp_path / "ref_main"
    r = subprocess.run(["g++", "-std=c++17", "-O0", f"-I{fp}", f"-I{os.path.join(ROOT, 'include')}",
                        str(src), *srcs, "-o", str(exe), f"-L{pkg}", "-lfemgpu",
                        f"-Wl,-rpath,{pkg}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
