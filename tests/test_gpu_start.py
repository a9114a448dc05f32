"""A model start from files stepped on the GPU: the bathymetry .a file and
a restart pair are read, masks and derived fields formed as the model's
start forms them (hycom_grid_masks, hycom_state_start), and the state is
stepped by hycom_hybgen_device against the oracle, bitwise."""
import numpy as np
import pytest

import hybgen_abi as abi
import oracle_lib as ol
from test_start import load_restart_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("hybmap", [2, 3])
def test_restart_state_on_device(gpu, tmp_path, hybmap):
    import torch
    print(f"\n[case] restart state 150x100x32 hybmap={hybmap} seed=301", flush=True)
    P, w = abi.workload_params("cfg1", hybmap=hybmap)
    P.ii, P.jj = P.idm, P.jdm
    d0 = abi.HostDomain(P).synth(abi.make_synth(P, w, seed=301))
    s = load_restart_state(tmp_path, P, d0)
    want = s.copy()
    ol.run_oracle(want, threads=8)
    dd = abi.DeviceDomain(P).upload(s)
    dd.hybgen(m=1, n=2)
    torch.cuda.synchronize()
    got = dd.download()
    df = ol.diff_domains(want, got)
    assert not df, f"arrays differ bitwise: {df}"
    assert not np.array_equal(want.a["dp"], s.a["dp"])
