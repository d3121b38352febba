"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and rejects bad arguments before touching a GPU."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "blink_mi355x.h")


def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"BLINK_API\s+[^;(]*?\b(blink_\w+)\s*\(", text)))


def test_header_declares_the_binding_symbols():
    from imager_amd import capi
    assert _declared_symbols() == sorted(capi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    from imager_amd import capi
    L = capi.load()
    for name in _declared_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (blink_\w+)", out))
    assert set(_declared_symbols()) <= exported
    # nothing else leaks with a blink_ prefix
    assert exported == set(_declared_symbols())


def test_library_targets_gfx950():
    from imager_amd import capi
    data = open(capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_version_and_fft_sizes():
    from imager_amd import capi
    assert "gfx950" in capi.version()
    for n in (2, 7, 8, 11, 13, 15, 128, 154, 180, 208, 256, 512, 1024, 2048, 4096, 6000, 8192):
        assert capi.fft_supported(n), n
    for n in (0, 1, 17, 34, 4097, 8194, 16384):
        assert not capi.fft_supported(n), n


def test_create_without_gpu_fails_cleanly():
    import torch
    from imager_amd import capi
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.BlinkError):
        capi.Context(0)


def test_image_set_struct_matches_header():
    """The ctypes mirror of blink_image_set has the C layout (size and offsets
    computed by the C compiler)."""
    import ctypes
    from imager_amd import capi
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "blink_mi355x.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(blink_image_set),
         offsetof(blink_image_set, corrections), offsetof(blink_image_set, delta_u),
         offsetof(blink_image_set, images), offsetof(blink_image_set, count_sums),
         offsetof(blink_image_set, transform), offsetof(blink_image_set, weighting),
         offsetof(blink_image_set, average));
  return 0;
}
'''
    import tempfile
    tmp = os.path.join(tempfile.mkdtemp(prefix="blink_abi_"), "abi_probe")
    with open(tmp + ".c", "w") as f:
        f.write(src)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", tmp, tmp + ".c"], check=True)
    vals = list(map(int, subprocess.run([tmp], capture_output=True, text=True,
                                        check=True).stdout.split()))
    S = capi.ImageSet
    assert vals == [ctypes.sizeof(S), S.corrections.offset, S.delta_u.offset, S.images.offset,
                    S.count_sums.offset, S.transform.offset, S.weighting.offset,
                    S.average.offset]
