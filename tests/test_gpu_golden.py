"""The HIP engine against the golden fixtures written from the compiled
reference core: every case, both grid backends, bit for bit (full arrays or
SHA-256 digests), plus per-atom counts from the count-only entry point."""

import numpy as np
import pytest

from test_oracle import ARR, CASES, case_inputs, check_against_golden

pytestmark = pytest.mark.gpu


def _engine(pos, cell, inv, pbc, cutoff, types, backend):
    import torch
    from matscipy_neighbours_amd import neighbours as nb
    rc, nums = nb._resolve_cutoff(cutoff, types if types is not None
                                  else np.ones(len(pos), np.int32))
    b = nb._Build(torch.from_numpy(np.ascontiguousarray(pos)).to("cuda:0"), np.zeros(3),
                  np.ascontiguousarray(cell), np.ascontiguousarray(inv), pbc, rc, nums,
                  backend=backend)
    out = b.run(31, want_counts=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("name", CASES)
def test_engine_matches_golden(gpu, name):
    pos, cell, inv, pbc, cutoff, types = case_inputs(name)
    for backend in (0, 1):
        res = _engine(pos, cell, inv, pbc, cutoff, types, backend)
        check_against_golden(name, res)
        assert np.array_equal(res["c"], ARR[f"{name}/counts"]), name
