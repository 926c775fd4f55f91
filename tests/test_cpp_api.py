"""The C++ plan classes of include/dedisp.hpp (reference src/Plan.hpp,
src/tdd/TDDPlan.hpp, src/fdd/FDDGPUPlan.hpp, src/fdd/FDDCPUPlan.hpp,
src/dedisp/DedispPlan.hpp): a program naming every class compiles and links
against libdedisp.so without a GPU, and the CPU FDD plan, which this GPU
library does not provide, refuses construction with DEDISP_UNKNOWN_ERROR."""
import subprocess
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]

SRC = r'''
#include "dedisp.hpp"
#include <cstdio>
#include <type_traits>
static_assert(std::is_base_of<dedisp::Plan, dedisp::TDDPlan>::value, "TDDPlan is a Plan");
static_assert(std::is_base_of<dedisp::TDDPlan, dedisp::DedispPlan>::value, "DedispPlan runs the TDD kernels");
static_assert(std::is_base_of<dedisp::Plan, dedisp::FDDGPUPlan>::value, "FDDGPUPlan is a Plan");
static_assert(std::is_base_of<dedisp::Plan, dedisp::FDDCPUPlan>::value, "FDDCPUPlan is a Plan");
int main(int argc, char**) {
    if (argc > 1) {  // never taken: the GPU classes only need to link
        dedisp::TDDPlan t(64, 64e-6f, 1500.f, -1.f);
        dedisp::DedispPlan d(64, 64e-6f, 1500.f, -1.f);
        dedisp::FDDGPUPlan f(64, 64e-6f, 1500.f, -1.f);
        return (int)(t.get_channel_count() + d.get_channel_count() + f.get_channel_count());
    }
    try {
        dedisp::FDDCPUPlan p(64, 64e-6f, 1500.f, -1.f, 0);
        std::puts("constructed");
        return 1;
    } catch (const dedisp::Error& e) {
        std::printf("%d\n", (int)e.code());
        return e.code() == DEDISP_UNKNOWN_ERROR ? 0 : 2;
    }
}
'''


def test_cpp_classes_compile_link_and_cpu_plan_refuses(dd, tmp_path):
    lib_dir = Path(dd.__file__).parent
    assert (lib_dir / "libdedisp.so").exists(), "build the library first"
    src = tmp_path / "api.cpp"
    src.write_text(SRC)
    exe = tmp_path / "api"
    r = subprocess.run(["g++", "-std=c++17", f"-I{REPO / 'include'}", str(src), "-o", str(exe), f"-L{lib_dir}",
                        "-ldedisp", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.stdout, r.stderr)
