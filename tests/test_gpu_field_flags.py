"""oi_field_flags and the field encoding built on it, against the same
decisions taken with torch on the host: which channels hold a non-zero value,
which hold a non-zero imaginary part, and the encoding chosen from them."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _expected(f: torch.Tensor):
    cin = f.shape[-1]
    flat = f.reshape(-1, cin)
    nonzero = (flat != 0).any(0).tolist()
    imag = (flat.imag != 0).any(0).tolist() if torch.is_complex(f) else [False] * cin
    return nonzero, imag


def _flags(g: torch.Tensor, kind: int, cin: int) -> int:
    from opticalimaging_amd import _native as nat
    word = torch.zeros(1, dtype=torch.int32, device="cuda")
    nat.check(nat.load().oi_field_flags(nat.ptr(g), kind, g.numel() // cin, cin, nat.ptr(word),
                                        nat.stream()))
    return int(word.item()) & 0xFFFFFFFF


@pytest.mark.parametrize("cin", [1, 3, 5, 16])
@pytest.mark.parametrize("cplx", [True, False])
def test_field_flags_bits(gpu, cin, cplx):
    from opticalimaging_amd import _native as nat
    rng = np.random.default_rng(100 + cin)
    for trial in range(6):
        shape = (2, 37, 53, cin)
        f = np.zeros(shape, dtype=np.complex64 if cplx else np.float32)
        for c in range(cin):
            mode = rng.integers(0, 4)   # 0 zero, 1 one real value, 2 one imaginary value, 3 dense
            if mode == 1:
                idx = tuple(int(rng.integers(0, s)) for s in shape[:3])
                f[idx + (c,)] = -0.25
            elif mode == 2 and cplx:
                idx = tuple(int(rng.integers(0, s)) for s in shape[:3])
                f[idx + (c,)] = 1j * 3e-30
            elif mode == 3:
                f[..., c] = rng.normal(size=shape[:3])
        if trial == 0:
            f[0, 0, 0, 0] = -0.0            # negative zero counts as zero
        if trial == 1 and cin > 1:
            f[1, 36, 52, cin - 1] = np.nan  # NaN counts as non-zero
        t = torch.as_tensor(f).cuda()
        bits = _flags(t.contiguous(), nat.OI_FIELD_C64_CHLAST if cplx else nat.OI_FIELD_REAL, cin)
        nonzero, imag = _expected(t)
        for c in range(cin):
            assert bool(bits >> (2 * c) & 1) == nonzero[c], (trial, c)
            assert bool(bits >> (2 * c + 1) & 1) == imag[c], (trial, c)


def test_field_encoding_choices(gpu):
    from opticalimaging_amd import _native as nat, engine
    rng = np.random.default_rng(7)
    base = rng.random((24, 40)).astype(np.float32)
    # one real live channel at index 2 -> float32 mask of that channel
    f = np.zeros((24, 40, 3), dtype=np.complex64)
    f[..., 2] = base
    enc = engine.field_encoding(torch.as_tensor(f).cuda())
    tens, kind, cin, chans = enc
    assert kind == nat.OI_FIELD_REAL and cin == 1 and chans == [0] and enc.source_channels == [2]
    assert torch.equal(tens.cpu()[0], torch.as_tensor(base))
    # an imaginary part keeps the complex encoding
    f[3, 4, 2] += 1e-3j
    tens, kind, cin, chans = engine.field_encoding(torch.as_tensor(f).cuda())
    assert kind == nat.OI_FIELD_C64_CHLAST and cin == 3 and chans == [2]
    # two live channels, real float64 input
    r = np.zeros((24, 40, 3))
    r[..., 0], r[..., 1] = base, -base
    enc = engine.field_encoding(torch.as_tensor(r).cuda())
    assert enc[1] == nat.OI_FIELD_C64_CHLAST and enc[3] == [0, 1] and enc[0].dtype == torch.complex64
    # an all-zero field keeps channel 0
    enc = engine.field_encoding(torch.zeros((8, 8, 3), device="cuda"))
    assert enc[1] == nat.OI_FIELD_REAL and enc.source_channels == [0]


def test_field_encoding_flags_follow_in_place_edits(gpu):
    """The channel flags attached to a field tensor are reused on an unchanged
    tensor and dropped by an in-place edit, so an edited channel is seen."""
    from opticalimaging_amd import _native as nat, engine
    t = torch.zeros((16, 24, 3), dtype=torch.complex64, device="cuda")
    t[..., 1] = 1.0
    enc = engine.field_encoding(t)
    assert enc[1] == nat.OI_FIELD_REAL and enc.source_channels == [1]
    assert engine.attached(t, "_oi_field_flags") is not None
    assert engine.field_encoding(t).source_channels == [1]          # from the attached flags
    t[2, 3, 0] = 0.5                                                # channel 0 becomes live
    assert engine.attached(t, "_oi_field_flags") is None
    enc = engine.field_encoding(t)
    assert enc[1] == nat.OI_FIELD_C64_CHLAST and enc[3] == [0, 1]
    t[..., 1].imag.fill_(0.25)                                      # channel 1 gains an imaginary part
    enc = engine.field_encoding(t)
    assert enc[1] == nat.OI_FIELD_C64_CHLAST and enc[3] == [0, 1]
    t.zero_()
    enc = engine.field_encoding(t)
    assert enc[1] == nat.OI_FIELD_REAL and enc.source_channels == [0]
    # the host tensor a caller keeps carries the flags too; a numpy field has none to keep
    h = torch.zeros((8, 8, 2))
    h[..., 1] = 2.0
    assert engine.field_encoding(h).source_channels == [1]
    assert engine.attached(h, "_oi_field_flags") == ((False, True), (False, False))
