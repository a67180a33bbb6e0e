"""Host-side bookkeeping of bench.py (no GPU): the bond-update count of one step, the
bulk bonds whose saturation the protocol checks, and the number-sector comparison of the
number-conserving check."""
import numpy as np

import bench


def test_updates_per_step_and_bulk():
    # 2nd-order step on L=128: odd bonds (64), even bonds (63), odd bonds (64)
    assert bench.N_UPDATES == 64 + 63 + 64
    # bond b (1-based) joins sites b and b+1; its cap min(d^b, d^(L-b), chi) is chi for
    # 4 <= b <= L-4 at d=4, chi=256 (4^4 = 256)
    links = np.arange(bench.L + 1)[bench.BULK]
    caps = np.minimum(np.minimum(float(bench.D) ** links, float(bench.D) ** (bench.L - links)), bench.CHI)
    assert links[0] == 4 and links[-1] == bench.L - 4
    assert np.all(caps == bench.CHI)
    assert min(bench.D ** 3, bench.CHI) < bench.CHI  # bond 3 is not in the bulk


def labelled_link(values, labels, chi=8):
    lam = np.zeros(chi)
    lab = np.full(chi, 10000, np.int32)
    lam[:len(values)] = values
    lab[:len(labels)] = labels
    return lam, lab


def test_sector_spectra_err():
    lam, lab = labelled_link([0.8, 0.5, 0.3, 0.1], [2, 1, 2, 3])
    LL, L0 = lab[None], lam[None]
    assert bench.sector_spectra_err(LL, L0, LL, L0) == 0.0
    # the same values of each sector in another order: no difference
    lam2, lab2 = labelled_link([0.8, 0.3, 0.5, 0.1], [2, 2, 1, 3])
    assert bench.sector_spectra_err(lab2[None], lam2[None], LL, L0) == 0.0
    # two values exchanged between sectors 1 and 2: the difference is seen
    lam3, lab3 = labelled_link([0.8, 0.5, 0.3, 0.1], [2, 2, 1, 3])
    assert abs(bench.sector_spectra_err(lab3[None], lam3[None], LL, L0) - 0.2) < 1e-15
    # a state missing from one sector counts with its whole value
    lam4, lab4 = labelled_link([0.8, 0.5, 0.3], [2, 1, 2])
    assert abs(bench.sector_spectra_err(lab4[None], lam4[None], LL, L0) - 0.1) < 1e-15
