"""
CPU checks of the FP32C sliding pass's LDS layout (fb_pass_layout, the host's pass_geometry): for
every line length, half kernel, segment length, workgroup size and tile kind below, the ranges
the kernel k_fast_pass (fb_kernels.hip) reads and writes stay inside the workgroup's LDS and
keep the zero padding it relies on (no GPU needed).

The kernel's accesses, per field half (v lines at float T+1 + l*stride, w lines after the
`lines` v lines; stride = len + gap):
* slide of lane sid (segment start s = sid*seg < len): reads s-T-1 .. s+T+seg of its line
  (the last seg-1 of them feed discarded steps past the line end);
* whole-segment stores when full = (len > seg and gap - (T+1) >= over), over = seg - len % seg
  (mod seg): the crossing lane stores s .. s+seg-1, which must stay short of the next line's
  left padding (the gap's last T+1 floats); it then re-zeroes min(T+1, over) floats after
  the cells, inside the gap;
* the injecting pass zeroes T+1 + stride floats from each line's left padding;
* the column tiles read element (l, r) for l < lines, r < len.
The reserve after the w lines covers the slides' reads past the last line.
"""
import ctypes

import pytest

SEGS = [19, 21, 23, 25, 29, 33, 37, 39, 41, 45, 49]
LENS = [7, 30, 64, 100, 128, 256, 300, 512, 600, 1200, 1249, 2400, 3000, 4800, 9000, 12544]
TS = [0, 1, 3, 9, 15, 27, 36, 44, 64, 90, 127, 200]


def layout(L, T, seg, threads, columns):
    from fastbarnes import _lib
    lpl, lines, gap = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lds = ctypes.c_int64()
    rc = _lib.lib().fb_pass_layout(L, T, seg, threads, columns, ctypes.byref(lpl), ctypes.byref(lines),
                                   ctypes.byref(gap), ctypes.byref(lds))
    return rc, lpl.value, lines.value, gap.value, lds.value


def check_layout(L, T, seg, threads, lpl, lines, gap, lds):
    nf = lds // 4                      # floats of LDS
    pad = T + 1
    stride = L + gap
    need = -(-L // seg)
    assert lpl >= need and lpl & (lpl - 1) == 0 and lpl // 2 < need or lpl == 1
    assert 1 <= lines and lines * lpl <= threads // 2
    assert gap >= pad
    assert lds <= 160 * 1024 and lds % 256 == 0
    assert 4 * (pad + 2 * lines * stride + 64) <= lds
    over = (seg - L % seg) % seg
    full = L > seg and gap - pad >= over
    for half in (0, 1):
        for l in range(lines):
            d0 = pad + (half * lines + l) * stride     # first cell of the line
            nxt_pad = d0 + stride - pad                # the next line's left padding starts here
            for sid in range(lpl):
                s = sid * seg
                if s >= L:
                    break
                assert d0 + s - T - 1 >= 0
                assert d0 + s + T + seg < nf, 'slide reads past the LDS'
                n = min(seg, L - s)
                end = d0 + s + (seg if (full or n == seg) else n)
                assert end <= nxt_pad, 'store reaches the next line\'s left padding'
            if full:
                # the re-zeroed floats after the cells stay inside the line's gap
                assert L + min(pad, over) <= stride
            # injecting pass: zeroes pad + stride floats from the left padding (d0 - pad)
            assert d0 + stride <= nf
            # column tiles: element (l, r), r < L
            assert d0 + L - 1 < nf


@pytest.mark.parametrize('threads', [128, 256, 512])
@pytest.mark.parametrize('columns', [0, 1])
def test_layout_bounds(threads, columns):
    ok = refused = 0
    for L in LENS:
        for T in TS:
            for seg in SEGS:
                rc, lpl, lines, gap, lds = layout(L, T, seg, threads, columns)
                if rc != 0:
                    refused += 1          # line too long for the workgroup or the LDS
                    continue
                check_layout(L, T, seg, threads, lpl, lines, gap, lds)
                ok += 1
    assert ok > 1000 and refused > 0


def _slide_banks(stride, lpl, seg, lines):
    worst = 0
    n = lines * lpl
    for t0 in range(0, n, 32):
        bank = [0] * 32
        for t in range(t0, min(t0 + 32, n)):
            b = ((t // lpl) * stride + (t % lpl) * seg) % 32
            bank[b] += 1
            worst = max(worst, bank[b])
    return worst


@pytest.mark.parametrize('L,T,seg,columns,stride', [
    (2400, 27, 39, 0, 2446),   # 2D x lines of 2400 nodes, sigma/step = 32: one line per half-wave
    (1200, 27, 39, 1, 1237),   # their y lines
    (512, 9, 33, 0, 560),      # 3D 512x512x128 x lines: 16 lanes, stride = 33*16*odd (mod 32)
    (512, 9, 33, 1, 560),
    (256, 9, 33, 1, 280),      # 3D batch 256x256x64 y lines
    (128, 9, 33, 1, 148),      # 3D z lines of 128
    (64, 9, 33, 1, 78),        # 3D batch z lines of 64: 16 lines per half-wave
])
def test_layout_strides_of_measured_classes(L, T, seg, columns, stride):
    """The line strides of the measured workload classes (256-thread workgroups); lines that
    share a half-wave get a stride whose slide lanes fall on distinct LDS banks."""
    rc, lpl, lines, gap, lds = layout(L, T, seg, 256, columns)
    assert rc == 0
    assert L + gap == stride
    assert _slide_banks(stride, lpl, seg, lines) == 1


def test_layout_rejects_bad_arguments():
    from fastbarnes import _lib
    assert layout(2400, 27, 40, 256, 0)[0] != 0     # no instantiation for an even segment
    assert layout(2400, 27, 39, 300, 0)[0] != 0     # workgroup size
    assert layout(0, 27, 39, 256, 0)[0] != 0
    assert layout(100000, 27, 39, 256, 0)[0] != 0   # longer than 128 lanes x 39
    assert _lib.lib().fb_last_error()
