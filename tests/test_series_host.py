"""Host side of the GPU series (mrsim_amd.series): the object-phase mirror against the oracle's
restatement of phase_slice / PrepPhase::shot (bit for bit: both are FP64 closed forms in the
reference's operation order), and the reference's refusals of acquire_volumes (src/kspace.rs:1863-1898),
which fire before any device work."""
import numpy as np
import pytest

from mrsim_amd import abi, series
from oracle import oracle as O


def _oracle_model(m: series.PhaseModel):
    om = O.PhaseModel()
    om.global_ = m.global_phase
    for i, c in enumerate(m.coeffs):
        om.coeffs[i] = c
    if m.prep is not None:
        om.has_prep, om.c_q, om.sigma_dx, om.sigma_rot = 1, m.prep.c_q, m.prep.sigma_dx, m.prep.sigma_rot
    return om


@pytest.mark.parametrize("snx,sny,o,z,nz", [(32, 32, 2, 1, 3), (17, 70, 1, 0, 1), (24, 27, 3, 8, 9)])
def test_phase_slice_and_shots_equal_the_oracle_bit_for_bit(snx, sny, o, z, nz):
    m = series.PhaseModel.hbcd_like()
    m.global_phase = 0.3
    m.coeffs = (0.1, 0.13, 0.10, 0.02, 1e-3, -2e-3, 5e-4, 1e-4, -3e-4, 2e-4)
    om = _oracle_model(m)
    for mag, d, g in [(1000.0, (1.0, 0.3, 0.0), 2), (3000.0, (0.0, 0.0, -2.0), 0), (0.0, (1.0, 0.0, 0.0), 1),
                      (800.0, (0.0, 0.0, 0.0), 3)]:
        s = series.prep_shot(m.prep, mag, d, g, z, 0xDEADBEEF12345)
        so = O.prep_shot(om, mag, d, g, z, 0xDEADBEEF12345)
        assert np.array_equal(s.view(np.uint64), so.view(np.uint64))
        p = series.phase_slice(m, s, snx, sny, o, z, nz)
        po = O.phase_slice(om, so, snx, sny, o, z, nz)
        assert np.array_equal(p.view(np.uint64), po.view(np.uint64))


def test_series_refusals():
    P = 32 * 32 * 3
    img = [np.ones(P * 2, dtype=np.float32)]
    acq = abi.Acquisition.default()
    ok = dict(sim_dims=(32, 32, 3), acq_dims=(16, 16, 3), n_volumes=2, images=img, t2=[80.0],
              fmap=np.zeros(P, dtype=np.float32), acq=acq)
    cases = [(dict(acq_dims=(16, 16, 2)), "slice count must match"),
             (dict(acq_dims=(15, 16, 3)), "integer multiple"),
             (dict(acq_dims=(16, 8, 3)), "oversampling must match"),
             (dict(t2=[80.0, 60.0]), "t2 has 2 entries for 1 compartments"),
             (dict(t_inhom=[40.0, 50.0]), "t_inhom has 2 entries"),
             (dict(eddy_drive=[None]), "eddy_drive has 1 entries for 2 volumes"),
             (dict(prep_drive=[None] * 3), "prep_drive has 3 entries for 2 volumes")]
    for over, msg in cases:
        with pytest.raises(ValueError, match=msg):
            series.SeriesAcquisition(**dict(ok, **over))
