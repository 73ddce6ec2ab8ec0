"""AddressSanitizer + UndefinedBehaviorSanitizer build of the host C++ (CPU
only): the library's input preparation, drop-in run glue and cell
methods/adjacency (rpd_prep.cpp, rpd_compat.cpp, rpd_cellhost.cpp) and the
oracle (rpd_oracle.cpp,
adj_oracle.cpp), compiled from source with g++ and driven by
tests/cpp/san_driver.cpp on a small clustered-sphere instance."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = ["tests/cpp/san_driver.cpp", "libmat_amd/csrc/host/rpd_prep.cpp",
           "libmat_amd/csrc/host/rpd_compat.cpp",
           "libmat_amd/csrc/host/rpd_cellhost.cpp", "oracle/rpd_oracle.cpp",
           "oracle/adj_oracle.cpp"]
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
         "-ffp-contract=off", "-pthread", "-I" + os.path.join(ROOT, "include")]


def test_host_code_under_asan_and_ubsan(tmp_path):
    objs, procs = [], []
    for src in SOURCES:
        obj = tmp_path / (os.path.basename(src) + ".o")
        objs.append(str(obj))
        procs.append(subprocess.Popen(
            ["g++"] + FLAGS + ["-c", os.path.join(ROOT, src), "-o", str(obj)],
            stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out = p.communicate(timeout=600)[0]
        assert p.returncode == 0, out.decode()[-2000:]
    exe = tmp_path / "san_driver"
    subprocess.run(["g++"] + FLAGS + objs + ["-o", str(exe)], check=True,
                   timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env,
                       timeout=300)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0 and "san ok" in r.stdout
    assert "runtime error" not in r.stderr
    assert "AddressSanitizer" not in r.stderr
