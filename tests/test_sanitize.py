"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU):
the mesh readers and geometry builders (prandtl_amd/csrc/mesh.cpp) and the
oracle, built from their sources by tests/sanitize/Makefile (g++), run on
every mesh file the tests hold, on malformed files (which must be refused
with PRANDTL_ERR_IO and a message, not crash or over-allocate) and on small
oracle RHS/RK3 runs.  A sanitizer report makes the test program exit
non-zero."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
MESHES = os.path.join(HERE, "golden", "meshes")

MALFORMED = {
    "truncated.mesh": "MFEM mesh v1.0\n\ndimension\n2\n\nelements\n3\n1 3 0 1 2 3\n1 3 9 9\n",
    "huge_count.mesh": "MFEM mesh v1.0\n\ndimension\n2\n\nelements\n2000000000\n1 3 0 1 2 3\n",
    "negative_count.mesh": "MFEM mesh v1.0\n\ndimension\n2\n\nelements\n-5\n",
    "bad_vertex_index.mesh": ("MFEM mesh v1.0\n\ndimension\n2\n\nelements\n1\n1 3 0 1 2 7\n\nboundary\n0\n\n"
                              "vertices\n4\n2\n0 0\n1 0\n1 1\n0 1\n"),
    "bad_nodes.msh": ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n2\n1 0 0 0\n$EndNodes\n"
                      "$Elements\n1\n1 3 2 0 0 1 2 7 99\n$EndElements\n"),
    "bad_element_node.msh": ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n"
                             "4 0 1 0\n$EndNodes\n$Elements\n1\n1 3 2 0 0 1 2 3 99\n$EndElements\n"),
    "huge_nodes.msh": "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1500000000\n1 0 0 0\n$EndNodes\n",
    "empty.mesh": "",
}


def test_host_code_under_sanitizers(tmp_path):
    subprocess.run(["make", "-C", SAN], check=True, capture_output=True, text=True)
    files = [os.path.join(MESHES, f) for f in sorted(os.listdir(MESHES)) if f.endswith((".mesh", ".msh"))]
    for name, text in MALFORMED.items():
        p = tmp_path / name
        p.write_text(text)
        files.append(str(p))
    garbage = tmp_path / "garbage.mesh"
    garbage.write_bytes(b"garbage\n\x00\xff\xfe" * 10)
    files.append(str(garbage))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(SAN, "_build", "san_driver")] + files + ["--oracle"],
                       capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    print(out[-4000:])
    assert r.returncode == 0, out[-4000:]
    assert r.stdout.rstrip().endswith("OK")
    for name in list(MALFORMED) + ["garbage.mesh"]:
        assert ("%s: refused (6)" % name) in r.stdout, name
    assert "Sanitizer" not in out and "runtime error" not in out
