"""GPU: the allocation-free entry point clmi_field_stats_noalloc against clmi_field_stats.

Both run the same kernels (the per-workgroup sums of f, f^2 and the forward-difference gradient
squared, then one fixed-order row sum); the allocation-free one keeps its partial rows in the
library's static device scratch and takes the field as a caller-owned device array.  Two calls in a
row, the second on a smaller lattice, must each equal the context entry point bit for bit, and sizes
outside the supported range or null pointers are refused before anything is launched.
"""
import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def test_field_stats_noalloc_matches_context_entry_point(gpu):
    import torch

    spec = gpu.lphi4(100.0, n_fields=1, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    rng = np.random.default_rng(3)
    stream = torch.cuda.current_stream()
    for N, dx in ((64, 0.4909), (32, 0.3)):
        say(f"case noalloc field stats N={N}")
        x = np.arange(N)
        f = 0.7 + 0.2 * np.cos(2 * np.pi * (3 * x[:, None, None] + x[None, :, None] - 2 * x[None, None, :]) / N)
        f = f + 0.05 * rng.normal(size=(N, N, N))
        lat = gpu.Lattice(spec, N, dx, 0.005, gpu.CLMI_F64)
        lat.upload(gpu.FLD_S, 0, 0, f)
        ref = lat.field_stats(gpu.FLD_S, 0, 0)
        lat.close()
        d = torch.from_numpy(f).to("cuda")
        out = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
        gpu.field_stats_noalloc(d.data_ptr(), N, dx, out.data_ptr(), stream.cuda_stream)
        stream.synchronize()
        got = tuple(float(v) for v in out.cpu().numpy())
        say("noalloc", got, "context", ref)
        assert got == ref
        # the mean and mean square equal numpy's to rounding
        assert got[0] == pytest.approx(float(np.mean(f)), rel=1e-12)
        assert got[1] == pytest.approx(float(np.mean(f * f)), rel=1e-12)
    # refused before any launch: a lattice beyond the supported range, a null pointer
    with pytest.raises(gpu.ClmiError, match="outside"):
        gpu.field_stats_noalloc(d.data_ptr(), 2048, 0.1, out.data_ptr(), stream.cuda_stream)
    with pytest.raises(gpu.ClmiError, match="null"):
        gpu.field_stats_noalloc(0, 32, 0.1, out.data_ptr(), stream.cuda_stream)
