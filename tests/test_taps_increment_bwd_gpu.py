"""Backward of the tap-increment layer (na_taps_increment_bwd) in each launch form
(split: the recomputed dz in a work buffer and an elementwise pass; split on the dz
the forward kept; fused: the weights-stationary product's epilogue, no workspace) against a float64
restatement of its formulas on the same bf16 operands:
  dz = D W^T, z = the point's centre pre-activation, dds_t = a + b_t (a the mean of
  the point's four rows), sp = softplus_beta,
  tap row  w (b_t sp'(z + dz) + a (sp'(z + dz) - sp'(z)))
  shared   a w sp'(z)
  centre  += sum_t dds_t w (sp'(z + dz_t) - sp'(z))
  db0     += column sums of the centre increments
  dws     += column sums of dds_t (sp(z + dz_t) - sp(z))."""
import ctypes

import pytest
import torch

from neuralangelo_amd import _lib

pytestmark = pytest.mark.gpu

BETA = 100.0


def _inputs(P, hid, K, seed):
    g = torch.Generator().manual_seed(seed)
    P32 = (P + 31) // 32 * 32
    D = torch.zeros(4 * P32, K)
    D[:4 * P] = torch.randn(4 * P, K, generator=g) * 1e-3
    W = torch.randn(hid, K, generator=g) * 0.05
    zc = torch.zeros(P32, hid)
    zc[:P] = torch.randn(P, hid, generator=g) * 0.02
    ws = torch.randn(hid, generator=g) * 0.05
    dds = torch.zeros(P32, 4)
    a = torch.randn(P, 1, generator=g) * 1e3
    dds[:P] = a + torch.randn(P, 4, generator=g) * 30.0
    centre = torch.zeros(P32, hid)
    centre[:P] = torch.randn(P, hid, generator=g) * 0.1
    return (D.to(torch.bfloat16), W.to(torch.bfloat16), zc, ws, dds,
            centre.to(torch.bfloat16), P32)


def _reference(D, W, zc, ws, dds, centre):
    P32, hid = zc.shape
    dz = (D.double() @ W.double().t()).view(P32, 4, hid)
    z = zc.double()[:, None, :]
    sig = lambda x: torch.sigmoid(BETA * x)  # noqa: E731
    sp = lambda x: torch.nn.functional.softplus(x, beta=BETA, threshold=1e9)  # noqa: E731
    d = dds.double()[:, :, None]
    a = d.mean(1, keepdim=True)
    s1, s0 = sig(z + dz), sig(z)
    w = ws.double()
    tap = w * ((d - a) * s1 + a * (s1 - s0))
    q = (a * w * s0)[:, 0]
    inc = (d * w * (s1 - s0)).sum(1)
    hv = sp(z + dz) - sp(z)
    return (tap.reshape(4 * P32, hid), q, centre.double() + inc, inc.sum(0),
            (d * hv).sum((0, 1)))


@pytest.mark.parametrize("mode", [0, 1, "kept"])
@pytest.mark.parametrize("P,hid,K", [(2000, 256, 288), (1000, 128, 160)])
def test_taps_increment_bwd_forms(cuda, mode, P, hid, K):
    print("taps increment bwd", mode, P, hid, K, flush=True)
    D, W, zc, ws, dds, centre, P32 = _inputs(P, hid, K, 5 + (mode if mode != "kept" else 2))
    ref = _reference(D, W, zc, ws, dds, centre)
    old = _lib.call("na_tune_set", b"taps_increment_bwd.fused", 1 if mode == 1 else 0)
    try:
        Dd, Wd, zd, wsd, ddsd = D.to(cuda), W.to(cuda), zc.to(cuda), ws.to(cuda), dds.to(cuda)
        cen = centre.to(cuda)
        dpre = torch.full((4 * P32, hid), 3.0, dtype=torch.bfloat16, device=cuda)
        q = torch.full((P32, hid), 3.0, dtype=torch.bfloat16, device=cuda)
        db0 = torch.zeros(hid, device=cuda)
        dws = torch.zeros(hid, device=cuda)
        work = torch.empty(4 * P32, hid, device=cuda) if mode != 1 else None
        held = 0
        if mode == "kept":
            # the forward keeps dz = D W^T in work; the backward then reads it
            parts = torch.empty(hid // 32, 4 * P32, device=cuda)
            _lib.call("na_taps_increment_fwd", _lib.ptr(Dd), K, 4 * P32, K, _lib.ptr(Wd), K, hid,
                      _lib.ptr(zd), hid, _lib.ptr(wsd), BETA, _lib.ptr(parts), _lib.ptr(work),
                      _lib.stream_of(Dd))
            torch.cuda.synchronize()
            dz_ref = D.double() @ W.double().t()
            scale = (D.double().abs() @ W.double().abs().t()).max().item()
            assert (work.cpu().double() - dz_ref).abs().max().item() <= 1e-5 * scale
            held = 1
        _lib.call("na_taps_increment_bwd", _lib.ptr(Dd), K, 4 * P32, K, _lib.ptr(Wd), K, hid,
                  _lib.ptr(zd), hid, _lib.ptr(wsd), BETA, _lib.ptr(ddsd), _lib.ptr(dpre), hid,
                  _lib.ptr(cen), hid, _lib.ptr(q), hid, _lib.ptr(db0), _lib.ptr(dws),
                  _lib.ptr(work), held, _lib.stream_of(Dd))
        torch.cuda.synchronize()
    finally:
        _lib.call("na_tune_set", b"taps_increment_bwd.fused", old)
    tap_r, q_r, cen_r, db0_r, dws_r = ref

    def close(name, got, want, rel):
        got = got.double().cpu()
        err = (got - want).abs()
        bound = rel * want.abs() + 1e-6 * want.abs().max()
        worst = (err / bound).max().item()
        print(f"  {name}: max |err| {err.max().item():.3e} ({worst:.2f} of the bound)")
        assert worst <= 1.0, name

    # bf16 outputs: one rounding (unit roundoff 2^-8) plus the float32 arithmetic
    # before it
    close("tap rows", dpre, tap_r, 2.0 ** -7)
    close("shared rows", q, q_r, 2.0 ** -7)
    close("centre rows", cen, cen_r, 2.0 ** -7)
    close("db0", db0, db0_r, 1e-4)
    close("dws", dws, dws_r, 1e-4)
