"""The C ABI: libcunfft_amd.so loads and exports every function that
include/cunfft.h declares (no compute call, so it runs without a GPU).

The library is built first when it is missing (hipcc cross-compiles gfx950
code on a machine without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("cunfft.h", "cunfft_util.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)          # comments
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)            # preprocessor lines
        names += re.findall(r"^\s*(?:const\s+)?(?:void|int|char|double|uint_t|GPUTimer)\s*\**\s*(\w+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


def load_library():
    import cunfft_amd

    if not os.path.exists(cunfft_amd.LIB_PATH):
        from cunfft_amd import build

        build.build(force=False)
    return ctypes.CDLL(cunfft_amd.LIB_PATH)


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    # the reference's plan surface (src/cunfft.h:35-261) ...
    for ref in ("cunfft_init", "cunfft_initGuru", "copyDataToDevice", "copyDataToDeviceAd",
                "cunfft_transform", "cunfft_adjoint", "copyDataToHost", "copyDataToHostAd",
                "copy_f_ToHost", "copy_f_hat_ToHost", "copy_g_ToHost", "copy_x_toHost",
                "cunfft_reinit", "cunfft_reinitAd", "cunfft_finalize", "cundft_transform",
                "cundft_adjoint", "resetDevice", "warmUp_createPlan", "cunfft_initHelp",
                "cunfft_initDevicePara", "cunfft_finalizeDevice", "cunfft_initPlan",
                "cunfft_freePlan", "resetCUNFFTTimes", "setProperties", "nfft_init_fg_exp_l"):
        assert ref in names, ref
    # ... and the additions: runtime m, errors, device pointers, gradients, FP32 family
    for add in ("cunfft_initGuru_m", "cunfft_last_error", "cunfft_transform_dev",
                "cunfft_adjoint_dev", "cunfft_backward_dev", "cunfftf_transform",
                "cunfftf_backward_dev"):
        assert add in names, add


def test_library_exports_every_declared_function():
    lib = load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
