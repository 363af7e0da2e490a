"""The reference's per-grain Python loops, run as batched GPU launches.

Each function reproduces one loop of ImageD11 over ``cImageD11`` calls —
call order, initial drlv2/labels arrays, label numbering — so its results
are those of the loop:

* `indexer_scores`      ``indexer.score`` over trial UBIs
                        (``ImageD11/indexing.py:818-820,1090-1097``)
* `fight_over_peaks`    ``indexer.fight_over_peaks``
                        (``ImageD11/indexing.py:876-905``): drlv2 = 2,
                        labels = -1, grain i gets label i
* `getind`              ``indexer.getind`` (``ImageD11/indexing.py:1060-1088``):
                        drlv2 = 1, labels = 0, label 1
* `assign_peaks_to_grains`  ``nb_utils.assign_peaks_to_grains``
                        (``ImageD11/nbGui/nb_utils.py:634-655``): drlv2 = 1,
                        labels = 0, grain i gets label i
* `assignlabels`        ``refinegrains.assignlabels``
                        (``ImageD11/refinegrains.py:684-734``): lab-frame
                        positions once, then compute_gv with each grain's
                        translation + score_and_assign, labels = -1,
                        drlv2 = 1
* `refine_grains`       ``refinegrains.refine`` (``ImageD11/refinegrains.py:
                        459-498``, triclinic): score_and_refine twice per
                        grain on its own g-vectors

All arrays are GPU tensors; nothing falls back to the CPU.
"""
import torch

from . import batched, geometry, refine


def _f64(x, dev):
    return torch.as_tensor(x, dtype=torch.float64, device=dev).contiguous()


def indexer_scores(ubis, gv, tol):
    """Peak counts of every trial UBI (int32 (K,))."""
    return batched.score_many(ubis, gv, tol)


def fight_over_peaks(ubis, gv, hkl_tol):
    """Returns (labels int32 (N,), drlv2 (N,), gas int32 (K,)) where gas[i]
    is the number of peaks finally labelled i (the histogram the reference
    stores in indexer.gas)."""
    N = gv.shape[0]
    drlv2 = torch.full((N,), 2.0, dtype=torch.float64, device=gv.device)
    labels = torch.full((N,), -1, dtype=torch.int32, device=gv.device)
    batched.assign_many(ubis, gv, hkl_tol, drlv2, labels)
    K = ubis.shape[0]
    lab = labels.long()
    gas = torch.bincount(lab[lab >= 0], minlength=K)[:K].to(torch.int32)
    return labels, drlv2, gas


def getind(ubi, gv, tol):
    """Boolean mask of the peaks indexed by one UBI."""
    N = gv.shape[0]
    drlv2 = torch.ones(N, dtype=torch.float64, device=gv.device)
    labels = torch.zeros(N, dtype=torch.int32, device=gv.device)
    one = torch.ones(1, dtype=torch.int32, device=gv.device)
    batched.assign_many(ubi.reshape(1, 3, 3).contiguous(), gv, tol, drlv2, labels,
                        grain_labels=one)
    return labels == 1


def assign_peaks_to_grains(ubis, gv, tol):
    """Returns (labels, drlv2, counts) with labels starting at 0."""
    N = gv.shape[0]
    drlv2 = torch.ones(N, dtype=torch.float64, device=gv.device)
    labels = torch.zeros(N, dtype=torch.int32, device=gv.device)
    counts = batched.assign_many(ubis, gv, tol, drlv2, labels)
    return labels, drlv2, counts


class Scan:
    """A peak table on the GPU with the per-peak quantities that do not
    depend on grains: lab positions (after removing any per-peak sample
    y-translation dty from yl), sin/cos of omega and the wedge/chi matrix.

    parity=True takes sin/cos of omega from the host's glibc, as the
    reference's compute_gv does (``batched.sincos_omega``): then every
    g-vector, label and drlv2 computed from this scan is the reference's,
    bit for bit.  parity=False (the fast mode) uses the device sincos,
    which differs from glibc's in the last bit for a few tenths of a
    percent of angles; labels only differ where a pair sits within ~1e-15
    relative of the tolerance.  The wedge/chi matrix is glibc's in both."""

    def __init__(self, sc, fc, omega, pars, dty=None, parity=False):
        dev = sc.device
        self.pars = geometry.full_parameters(pars)
        self.parity = bool(parity)
        sign, wvln, wedge, chi = geometry.pack_gv(self.pars)
        self.mat = batched.wedge_chi_matrix(wedge, chi)
        self.gv0, self.xl, self.so, self.co = batched.sf2gv(sc, fc, omega, self.pars, keep=True,
                                                            parity=parity)
        if dty is not None:
            self.xl[:, 1] -= dty
            self.gv0 = batched.gvectors_trig(self.xl, self.so, self.co, self.mat, wvln)
        self.n = sc.shape[0]
        self.device = dev


def assignlabels(scan, ubis, translations, tol, grain_labels=None):
    """refinegrains.assignlabels: returns (labels, drlv2, counts)."""
    dev = scan.device
    labels = torch.full((scan.n,), -1, dtype=torch.int32, device=dev)
    drlv2 = torch.ones(scan.n, dtype=torch.float64, device=dev)
    sign, wvln, wedge, chi = geometry.pack_gv(scan.pars)
    counts = batched.gv_assign_many(scan.xl, scan.so, scan.co, wvln, ubis, translations, tol,
                                    drlv2, labels, grain_labels=grain_labels, mat=scan.mat)
    return labels, drlv2, counts


def grain_gvectors(scan, translations, labels, K):
    """The g-vectors of each grain's assigned peaks, computed with that
    grain's translation: (gv_lists (M,3), idx (M,), off (K+1,))."""
    idx, off, owner = batched.lists_from_labels(labels, K)
    sign, wvln, wedge, chi = geometry.pack_gv(scan.pars)
    gvl = batched.gv_lists(scan.xl, scan.so, scan.co, wvln, translations, idx, owner,
                           mat=scan.mat)
    return gvl, idx, off


def refine_grains(ubis, gv_lists, off, tol, n_iter=2, config=None):
    """n_iter score_and_refine calls per grain on each grain's own list of
    g-vectors (gv_lists rows off[g]..off[g+1]).  Returns (ubis, n, mean)."""
    if config is None:
        return batched.refine_lists(ubis, gv_lists, None, off, tol, n_iter)
    return refine.refine(ubis, gv_lists, None, off, tol, n_iter, config=config)
