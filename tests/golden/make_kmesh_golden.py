"""Write the .k -> prepared-mesh golden vectors of tests/golden/kmesh/.

Runs the reference's mesh parser (mesh_parser_lib/parser_src/parser_lib.py,
FileParser; the call sequence of mesh_parser_lib/fem_parser.py:4-22) on .k
meshes of the reference's raw_meshes/ whose parser branches no committed
fixture pins (3D segments and node-set loads), in this container only, and
stores each input .k and the files it writes.  Usage:

    python tests/golden/make_kmesh_golden.py /root/reference
"""
import contextlib
import io
import os
import shutil
import sys
import tempfile

CASES = [("3D", "small_test"), ("3D", "bulk_small"), ("2D", "test_rect_pcg")]
FILES = ["nodes.txt", "elements.txt", "constraints.txt", "loads.txt", "stress.txt"]


def main(ref):
    sys.path.insert(0, os.path.join(ref, "mesh_parser_lib", "parser_src"))
    from parser_lib import FileParser  # the reference parser, read-only
    out_root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kmesh")
    for dim, name in CASES:
        raw = os.path.join(ref, "raw_meshes", dim, name + ".k")
        out = os.path.join(out_root, name)
        os.makedirs(out, exist_ok=True)
        shutil.copy(raw, os.path.join(out, name + ".k"))
        with tempfile.TemporaryDirectory() as tmp:
            p = FileParser(name + ".k", tmp, raw, name, int(dim[0]))
            with contextlib.redirect_stdout(io.StringIO()):
                p.parse_nodes()
                p.prepare_nodes()
                p.parse_loads()
                p.parse_elements()
                p.prepare_elements()
                p.parse_constraints_and_sets()
                p.prepare_constraints()
                p.prepare_loads()
            for f in FILES:
                shutil.copy(os.path.join(tmp, f), os.path.join(out, "ref_" + f))
        print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
