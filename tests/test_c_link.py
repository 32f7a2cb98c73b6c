"""Link check of the C ABI as a C caller sees it: a C11 program includes
include/wgpu_nufft.h and include/mnufft.h, links against libmnufft.so and
runs the calls that need no GPU (options layout and defaults, version,
host-side planning queries, error codes of plan creation on a host without
a GPU).  The reference's Fortran and Julia wrappers bind exactly these
prototypes (capi/include/wgpu_nufft.h:89-269)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <assert.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "wgpu_nufft.h"
#include "mnufft.h"

int main(void)
{
    _Static_assert(sizeof(wgpu_nufft_opts) == 320, "opts layout");
    _Static_assert(offsetof(wgpu_nufft_opts, sigma) == 16, "sigma offset");
    wgpu_nufft_opts o;
    memset(&o, 0xff, sizeof o);
    wgpu_nufft_default_opts(&o);
    assert(o.backend == WGPU_NUFFT_BACKEND_AUTO && o.precision == WGPU_NUFFT_PRECISION_AUTO);
    assert(o.sigma == 0.0 && o.adapter_index == 0 && o.adapter_name[0] == 0);
    assert(strcmp(wgpu_nufft_version(), "0.3.2") == 0);
    int w = 0;
    double beta = 0.0;
    assert(mnufft_kernel_params(1e-6, 2.0, &w, &beta) == MNUFFT_SUCCESS);
    assert(w == 7 && beta > 16.0 && beta < 16.2);
    if (getenv("MNUFFT_LINK_NO_GPU")) {
        wgpu_nufftf_plan plan = NULL;
        int64_t modes[2] = {8, 8};
        int rc = wgpu_nufftf_makeplan(1, 2, modes, 1, 1, 1e-6, &plan, NULL);
        assert(rc == WGPU_NUFFT_ERROR_GPU_UNAVAILABLE && plan == NULL);
        assert(strlen(wgpu_nufft_last_error()) > 0);
    }
    puts("c link ok");
    return 0;
}
"""


def test_c_program_links_and_runs(tmp_path):
    import torch
    lib_dir = os.path.join(ROOT, "wgpunufft_amd", "_lib")
    if not os.path.exists(os.path.join(lib_dir, "libmnufft.so")):
        pytest.fail("libmnufft.so is not built")
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "link.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "link"
    subprocess.run([gcc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", lib_dir, "-lmnufft",
                    "-Wl,-rpath," + lib_dir], check=True)
    env = dict(os.environ)
    if not torch.cuda.is_available():
        env["MNUFFT_LINK_NO_GPU"] = "1"
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, env=env)
    assert "c link ok" in out.stdout
