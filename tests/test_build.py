"""The C++ mirror header compiles against the C ABI, and the app built over
it links (CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mirror_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "graphit_amd.hpp"\nint main(){graphit::SimpleGPUSchedule s;'
                   's.configLoadBalancing(graphit::SimpleGPUSchedule::load_balancing_type::TWCE);'
                   'return s.s.load_balancing == GI_LB_TWCE ? 0 : 1;}\n')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", os.path.join(ROOT, "easy_graphit_amd"), "-lgraphit_amd",
                    f"-Wl,-rpath,{os.path.join(ROOT, 'easy_graphit_amd')}"], check=True)
    assert subprocess.run([str(exe)]).returncode == 0
