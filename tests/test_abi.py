"""The C ABI library: loads, exports every declared entry, layouts match C."""
import os
import re
import subprocess

import pytest

import lyapunov3d_amd as lyap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("lyap.h", "lyap_types.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lyap_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = lyap.load()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(lyap.exported_symbols())


def test_no_device_is_an_error_not_a_crash():
    import ctypes as C
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    rc = lyap.load().lyap_create(C.byref(h), 0, 0)
    assert rc == lyap.ERR_NO_DEVICE and not h.value


def test_strerror():
    assert lyap.strerror(lyap.ERR_BAD_SEQUENCE) == "bad sequence"
    assert lyap.strerror(12345) == "unknown status"


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "lyap.h"
#define S(T) printf(#T " %zu %zu\n", sizeof(T), _Alignof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(LyapParams) S(LyapPoint) S(LyapCam) S(LyapRGBA) S(LyapView) S(LyapRequest)
  S(LyapVolumeMode) S(LyapRenderMode) S(LyapTuning) S(LyapServerConfig) S(LyapResult) S(LyapServerStats)
  S(LyapEditor) S(LyapVolumeBlock)
  O(LyapTuning, level_budget) O(LyapTuning, voxel_tile) O(LyapVolumeBlock, seq) O(LyapVolumeBlock, n)
  O(LyapVolumeBlock, out)
  O(LyapCam, Q) O(LyapCam, ambient) O(LyapCam, chaosColor) O(LyapView, prm) O(LyapView, seq)
  O(LyapView, rgba) O(LyapView, num_lights) O(LyapRequest, view) O(LyapRequest, vol_prm)
  O(LyapRequest, vol_out) O(LyapRequest, deadline_ms) O(LyapServerConfig, mode) O(LyapServerConfig, vol_mode)
  O(LyapResult, submit_ms) O(LyapResult, lambda) O(LyapServerConfig, slice_ms)
  O(LyapServerConfig, max_queued)
  O(LyapServerStats, resident_bytes) O(LyapServerStats, device_ms)
  O(LyapEditor, cam) O(LyapEditor, lights) O(LyapEditor, home_light) O(LyapEditor, num_lights)
  O(LyapEditor, pending)
  return 0;
}
"""


def test_struct_layouts_match_c(tmp_path):
    import ctypes as C
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c = {}
    for line in out:
        if line.strip():
            parts = line.split()
            c[parts[0]] = [int(x) for x in parts[1:]]
    py = {"LyapParams": lyap.LyapParams, "LyapPoint": lyap.LyapPoint, "LyapCam": lyap.LyapCam,
          "LyapRGBA": lyap.RGBA, "LyapView": lyap.View, "LyapRequest": lyap.Request,
          "LyapVolumeMode": lyap.VolumeMode, "LyapRenderMode": lyap.RenderMode, "LyapTuning": lyap.Tuning,
          "LyapServerConfig": lyap.ServerConfig, "LyapResult": lyap.Result,
          "LyapServerStats": lyap.ServerStats, "LyapEditor": lyap.Editor, "LyapVolumeBlock": lyap.VolumeBlock}
    for name, cls in py.items():
        assert C.sizeof(cls) == c[name][0], name
    assert c["LyapParams"][0] == 56 and c["LyapPoint"][0] == 36 and c["LyapCam"][0] == 224
    fields = {"LyapCam.Q": lyap.LyapCam.Q, "LyapCam.ambient": lyap.LyapCam.ambient,
              "LyapCam.chaosColor": lyap.LyapCam.chaosColor, "LyapView.prm": lyap.View.prm,
              "LyapView.seq": lyap.View.seq, "LyapView.rgba": lyap.View.rgba,
              "LyapView.num_lights": lyap.View.num_lights, "LyapRequest.view": lyap.Request.view,
              "LyapRequest.vol_prm": lyap.Request.vol_prm, "LyapRequest.vol_out": lyap.Request.vol_out,
              "LyapRequest.deadline_ms": lyap.Request.deadline_ms, "LyapServerConfig.mode": lyap.ServerConfig.mode,
              "LyapServerConfig.vol_mode": lyap.ServerConfig.vol_mode, "LyapResult.submit_ms": lyap.Result.submit_ms,
              "LyapResult.lambda": lyap.Result.lambda_, "LyapServerConfig.slice_ms": lyap.ServerConfig.slice_ms,
              "LyapServerConfig.max_queued": lyap.ServerConfig.max_queued,
              "LyapServerStats.resident_bytes": lyap.ServerStats.resident_bytes,
              "LyapServerStats.device_ms": lyap.ServerStats.device_ms, "LyapEditor.cam": lyap.Editor.cam,
              "LyapEditor.lights": lyap.Editor.lights, "LyapEditor.home_light": lyap.Editor.home_light,
              "LyapEditor.num_lights": lyap.Editor.num_lights, "LyapEditor.pending": lyap.Editor.pending,
              "LyapTuning.level_budget": lyap.Tuning.level_budget, "LyapTuning.voxel_tile": lyap.Tuning.voxel_tile,
              "LyapVolumeBlock.seq": lyap.VolumeBlock.seq, "LyapVolumeBlock.n": lyap.VolumeBlock.n,
              "LyapVolumeBlock.out": lyap.VolumeBlock.out}
    for k, f in fields.items():
        assert f.offset == c[k][0], k


def test_cpp_driver_builds_and_reports_no_device():
    """The C++ driver (tests/driver) compiles against lyap.h and links
    liblyap_amd.so; without a GPU it exits with LYAP_ERR_NO_DEVICE's code."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    d = os.path.join(ROOT, "tests", "driver")
    subprocess.run(["make", "-C", d], check=True, capture_output=True)
    r = subprocess.run([os.path.join(d, "lyap_driver"), "calculate", "8", "plain", os.devnull],
                       capture_output=True, text=True)
    assert r.returncode == -lyap.ERR_NO_DEVICE, r.stderr
    assert "no device" in r.stderr
    r = subprocess.run([os.path.join(d, "lyap_driver")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
