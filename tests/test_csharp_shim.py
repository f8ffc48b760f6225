"""The C# drop-in class MultiGridSchemeHIP<T> published in INTEGRATION.md,
replayed in the reference model's call order.

The model writes the Dirichlet data and the initial guess into the scheme's
host mirror un0 through Iterative1DScheme<T>'s utilities, each of which
starts with getArray() (VisualLaplacePoisson2D/Model/CUDA/
Iterative1DScheme.cs:20-61), in the order top/bottom borders, left/right
borders, initial guess (VLPRectangleModel.cs:541-546), then uploads
(initAfterBoundariesAndInitialIterationInited) and iterates.  ShimMirror
restates the C# class statement by statement; the first test checks that the
C# text in INTEGRATION.md is exactly the one mirrored, the others replay the
call order on a fake device (CPU) and through the library (GPU) and check
that the values written before the upload are the ones uploaded."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle.setup import boundary, init_guess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class ShimMirror:
    """Statement-by-statement restatement of MultiGridSchemeHIP<T>'s
    ISchemeCalculate methods and getArray; CS holds
    the C# statements each Python method mirrors, in order."""

    CS = {
        "getArray": [
            "if (un0Changed)", "{", "un0Changed = false;",
            "NativeMG.Check(h, NativeMG.vlp2d_mg_download(h, un0d));",
            "for (int k = 0; k < un0.Length; k++) un0[k] = T.CreateTruncating(un0d[k]);",
            "}", "return un0;"],
        "initAfterBoundariesAndInitialIterationInited": [
            "for (int k = 0; k < un0.Length; k++) un0d[k] = double.CreateTruncating(un0[k]);",
            "NativeMG.Check(h, NativeMG.vlp2d_mg_upload(h, un0d));",
            "un0Changed = false;"],
        "doIteration": [
            "un0Changed = true;",
            "NativeMG.Check(h, NativeMG.vlp2d_mg_cycle(h, out double m));",
            "return m > double.CreateTruncating(eps) ? eps + eps : T.Zero;"],
    }

    def __init__(self, native, seg_x, seg_y, eps):
        self.nat = native
        self.dimX, self.dimY = seg_x + 1, seg_y + 1
        self.eps = eps
        self.un0 = np.zeros(self.dimX * self.dimY)
        self.un0d = np.zeros(self.dimX * self.dimY)
        self.un0Changed = False  # a C# bool field starts false
        self.h = native.create(seg_x, seg_y)

    def getArray(self):
        if self.un0Changed:
            self.un0Changed = False
            self.nat.check(self.nat.download(self.h, self.un0d))
            self.un0[:] = self.un0d
        return self.un0

    def initAfterBoundariesAndInitialIterationInited(self):
        self.un0d[:] = self.un0
        self.nat.check(self.nat.upload(self.h, self.un0d))
        self.un0Changed = False

    def doIteration(self, it):
        self.un0Changed = True
        rc, m = self.nat.cycle(self.h)
        self.nat.check(rc)
        return self.eps + self.eps if m > self.eps else 0.0


class PublishedBeforeFix(ShimMirror):
    """getArray as INTEGRATION.md published it before: an unconditional
    download (kept to show that the replay below detects that mistake)."""

    def getArray(self):
        self.nat.check(self.nat.download(self.h, self.un0d))
        self.un0[:] = self.un0d
        return self.un0


class Iterative1D:
    """The Iterative1DScheme<T> utilities the model calls before the upload
    (Iterative1DScheme.cs:20-61): each works on getArray()'s array in place."""

    def __init__(self, scheme, hx, hy, border):
        self.s, self.hx, self.hy, self.border = scheme, hx, hy, border

    def _grid(self):
        return self.s.getArray().reshape(self.s.dimX, self.s.dimY)

    def initTopBottomBorders(self):
        u, _, _ = boundary(self.s.dimX, self.s.dimY, self.hx, self.hy, border=self.border)
        a = self._grid()
        a[:, 0], a[:, -1] = u[:, 0], u[:, -1]

    def initLeftRightBorders(self):
        u, _, _ = boundary(self.s.dimX, self.s.dimY, self.hx, self.hy, border=self.border)
        a = self._grid()
        a[0, :], a[-1, :] = u[0, :], u[-1, :]

    def initInitialIterationMean(self, val):
        a = self._grid()
        a[1:-1, 1:-1] = val


class FakeNative:
    """A device on the host: one array per handle; a cycle halves the interior."""

    def __init__(self):
        self.dev, self.downloads = {}, 0

    def create(self, sx, sy):
        self.dev[1] = np.zeros((sx + 1) * (sy + 1))
        return 1

    def upload(self, h, a):
        self.dev[h][:] = a
        return 0

    def download(self, h, a):
        self.downloads += 1
        a[:] = self.dev[h]
        return 0

    def cycle(self, h):
        self.dev[h] *= 0.5
        return 0, 1.0

    @staticmethod
    def check(rc):
        assert rc == 0


class LibNative:
    """The solver library through its C ABI (vlp2d_mg_create/upload/download/cycle)."""

    def __init__(self, hx, hy):
        from vlp2d_amd import _lib
        self.L, self.lib = _lib, _lib.load_library()
        self.hx, self.hy, self.handles = hx, hy, []

    def create(self, sx, sy):
        h = ctypes.c_void_p()
        self.check(self.lib.vlp2d_mg_create(sx, sy, self.hx, self.hy, None, 1e-13, 0, None, ctypes.byref(h)))
        self.handles.append(h)
        return h

    def upload(self, h, a):
        return self.lib.vlp2d_mg_upload(h, self.L.ptr(a))

    def download(self, h, a):
        return self.lib.vlp2d_mg_download(h, self.L.ptr(a))

    def cycle(self, h):
        m = ctypes.c_double()
        rc = self.lib.vlp2d_mg_cycle(h, ctypes.byref(m))
        return rc, m.value

    def check(self, rc):
        assert rc == 0, self.lib.vlp2d_mg_last_error(None)

    def close(self):
        for h in self.handles:
            self.lib.vlp2d_mg_destroy(h)


def _csharp_method(text, name):
    """Statements of method `name` of the MultiGridSchemeHIP C# block."""
    m = re.search(r"```csharp\n(.*?class MultiGridSchemeHIP<T>.*?)```", text, re.S)
    assert m, "MultiGridSchemeHIP<T> code block not found"
    code = m.group(1)
    k = re.search(r"\b" + name + r"\s*\([^)]*\)\s*\n?\s*\{", code)
    assert k, name
    depth, i = 1, k.end()
    start = i
    while depth:
        depth += {"{": 1, "}": -1}.get(code[i], 0)
        i += 1
    body = code[start:i - 1]
    out = []
    for line in body.split("\n"):
        line = line.split("//")[0].strip()
        if line:
            out.append(re.sub(r"\s+", " ", line))
    return out


def test_published_csharp_is_the_mirrored_text():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for name, stmts in ShimMirror.CS.items():
        assert _csharp_method(text, name) == stmts, name
    # the flag the mirror relies on is a field of the class
    assert re.search(r"\bbool un0Changed;", text)


def _replay(cls, native, seg, h, val):
    s = cls(native, seg, seg, 1e-13)
    border = lambda x, y: x * x - y * y + 1.0
    it = Iterative1D(s, h, h, border)
    it.initTopBottomBorders()
    it.initLeftRightBorders()
    it.initInitialIterationMean(val)
    s.initAfterBoundariesAndInitialIterationInited()
    u, _, _ = boundary(seg + 1, seg + 1, h, h, border=border)
    return s, init_guess(u, 0, val).ravel()


def test_replay_keeps_the_borders_on_a_fake_device():
    nat = FakeNative()
    s, want = _replay(ShimMirror, nat, 16, 1.0 / 16, 0.25)
    assert np.array_equal(nat.dev[s.h], want)  # what was written before the upload is what was uploaded
    assert nat.downloads == 0
    s.doIteration(0)
    assert np.array_equal(s.getArray(), 0.5 * want) and nat.downloads == 1
    s.getArray()
    assert nat.downloads == 1  # no cycle since: un0 is current
    # the unconditional download published before loses the borders to the
    # device's zeros before the upload
    nat2 = FakeNative()
    s2, want2 = _replay(PublishedBeforeFix, nat2, 16, 1.0 / 16, 0.25)
    assert not np.array_equal(nat2.dev[s2.h], want2)
    assert np.all(nat2.dev[s2.h].reshape(17, 17)[:, 0] == 0.0)


@pytest.mark.gpu
def test_replay_keeps_the_borders_through_the_library():
    seg, h = 64, 1.0 / 64
    nat = LibNative(h, h)
    try:
        s, want = _replay(ShimMirror, nat, seg, h, 0.25)
        got = np.zeros_like(want)
        nat.check(nat.download(s.h, got))
        assert np.array_equal(got, want)
        s.doIteration(0)
        nat.check(nat.download(s.h, got))
        assert np.array_equal(s.getArray(), got)
        b = got.reshape(seg + 1, seg + 1)
        w = want.reshape(seg + 1, seg + 1)
        assert np.array_equal(b[0, :], w[0, :]) and np.array_equal(b[:, -1], w[:, -1])  # Dirichlet data kept
    finally:
        nat.close()
