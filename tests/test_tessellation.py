"""S^2 tessellation, conjugate augmentation and the RELION STAR reader (CPU): the oracle's
pins and the host parts of the product against them (S2tessellation.py, distribute3Sphere.py,
util.py:76-106, read_alignfile.py:20-96)."""
import os

import numpy as np
import pytest

from oracle import tessellation as ot

DEMO_STAR = "/root/reference/demo/RyR1GCs_clustRem.star"


def _random_q(n, seed):
    q = np.random.default_rng(seed).normal(size=(4, n))
    return q / np.linalg.norm(q, axis=0)


@pytest.mark.parametrize("K", [10, 100, 700, 2000])
def test_distribute3sphere_rings(K):
    """Points on the unit sphere, K of them, on rings of polar angle (k + 1/2) delta with
    equally spaced azimuths; the product's vectorised rings equal the loop restatement."""
    from manifoldem_python_amd import distribute3Sphere
    pts, it = ot.distribute3Sphere(K)
    assert pts.shape == (K, 3) and it <= 100
    np.testing.assert_allclose(np.linalg.norm(pts, axis=1), 1.0, atol=1e-14)
    w1 = np.unique(np.round(np.arccos(np.clip(pts[:, 0], -1, 1)), 10))
    if len(w1) > 2:
        d = np.diff(w1)
        np.testing.assert_allclose(d, d[0], rtol=1e-6)
        np.testing.assert_allclose(w1[0], d[0] / 2, rtol=1e-6)
    mine, it2 = distribute3Sphere.op(K)
    assert it2 == it
    np.testing.assert_allclose(mine, pts, rtol=0, atol=1e-15)


def test_augment_gives_antipodal_directions():
    """S2 of [-q1, q0, -q3, q2] is -S2(q) for unit q; S2 of the identity is the z axis."""
    from manifoldem_python_amd import util, S2tessellation
    q = _random_q(500, 1)
    qa = util.augment(q)
    np.testing.assert_array_equal(qa, ot.augment(q))
    S2 = ot.get_S2(qa)
    np.testing.assert_allclose(S2[:, 500:], -S2[:, :500], atol=1e-14)
    np.testing.assert_allclose(ot.get_S2(np.array([[1.0], [0], [0], [0]]))[:, 0], [0, 0, 1])
    np.testing.assert_array_equal(S2tessellation.get_S2(qa), S2)


def test_oracle_classS2_is_the_nearest_direction():
    rng = np.random.default_rng(2)
    X = ot.distribute3Sphere(300)[0]
    Q = ot.get_S2(_random_q(4000, 3)).T
    IND, NC = ot.classS2(X, Q)
    d = ((Q[:, None, :] - X[None]) ** 2).sum(-1)
    np.testing.assert_array_equal(IND[:, 0], np.argmin(d, axis=1))
    np.testing.assert_array_equal(NC, np.bincount(np.argmin(d, axis=1)))


def test_select_bins_both_halves():
    """The halving rule of S2tessellation.py:85-108 on hand-built occupancies: bincount stops
    at the last occupied bin, so a short NC keeps the first half; otherwise the second half."""
    from manifoldem_python_amd import S2tessellation
    NC = np.array([5, 0, 30, 2, 40, 1])                # 10 bins, NC of length 6: mid = 5
    assert ot.select_bins(NC, 10, 3) == [0, 2, 4]
    NC = np.array([50, 0, 4, 9, 0, 7, 3, 8, 1, 9])     # 10 bins, full NC: halves of 5 and 5
    assert ot.select_bins(NC, 10, 5) == [0, 3]         # len(NC1) >= len(NC2): first half
    NC = np.array([9, 9, 9, 9, 1, 9, 9])               # 7 bins: mid = 3, NC2 longer
    assert ot.select_bins(NC, 7, 9) == [3, 5, 6]
    rng = np.random.default_rng(4)
    for nb in (11, 40, 41):
        NC = rng.integers(0, 50, size=rng.integers(nb // 2, nb + 1))
        assert S2tessellation.select_bins(NC, nb, 25) == ot.select_bins(NC, nb, 25)


def _write_star31(path, rot, tilt, psi, U, V, ox, oy):
    with open(path, "w") as f:
        f.write("# version 30001\n\ndata_optics\n\nloop_\n_rlnOpticsGroupName #1\n"
                "_rlnOpticsGroup #2\n_rlnVoltage #3\n_rlnSphericalAberration #4\n"
                "_rlnAmplitudeContrast #5\n_rlnImagePixelSize #6\nopticsGroup1 1 200.0 "
                "2.7 0.07 1.1\n\n\n# version 30001\n\ndata_particles\n\nloop_\n"
                "_rlnImageName #1\n_rlnAngleRot #2\n_rlnAngleTilt #3\n_rlnAnglePsi #4\n"
                "_rlnDefocusU #5\n_rlnDefocusV #6\n_rlnOriginX #7\n_rlnOriginY #8\n")
        for i in range(len(rot)):
            f.write(f"{i + 1:06d}@stack.mrcs {rot[i]:.6f} {tilt[i]:.6f} {psi[i]:.6f} "
                    f"{U[i]:.6f} {V[i]:.6f} {ox[i]:.6f} {oy[i]:.6f}\n")
        f.write("\n")


def test_star31_reader(tmp_path):
    from manifoldem_python_amd import star
    rng = np.random.default_rng(5)
    n = 37
    rot, tilt, psi = rng.uniform(-180, 180, n), rng.uniform(0, 180, n), rng.uniform(-180, 180, n)
    U, V = rng.uniform(5000, 30000, n), rng.uniform(5000, 30000, n)
    ox, oy = rng.normal(size=n), rng.normal(size=n)
    path = os.path.join(tmp_path, "p.star")
    _write_star31(path, rot, tilt, psi, U, V, ox, oy)
    sh, q, U2, V2, micro = star.get_from_relion(path)
    assert micro == dict(EkV=200.0, Cs=2.7, AmpContrast=0.07)
    np.testing.assert_allclose(U2, U, atol=1e-6)
    np.testing.assert_allclose(sh[1], oy, atol=1e-6)
    ref = ot.euler_q(np.deg2rad(np.round(rot, 6)), np.deg2rad(np.round(tilt, 6)),
                     np.deg2rad(np.round(psi, 6)))
    np.testing.assert_allclose(q, ref, atol=1e-12)


@pytest.mark.skipif(not os.path.exists(DEMO_STAR), reason="reference demo data not present")
def test_demo_star_reads_and_tessellates_on_the_host_oracle():
    """The reference's demo alignment (RELION, no optics block): 14491 particles, its
    microscope columns, unit quaternions; the oracle tessellation of the augmented set puts
    each particle and its conjugate in antipodal bins' halves."""
    from manifoldem_python_amd import star, util
    sh, q, U, V, micro = star.get_from_relion(DEMO_STAR)
    assert q.shape == (4, 14491)
    assert micro == dict(EkV=300.0, Cs=2.26, AmpContrast=0.1)
    np.testing.assert_allclose(np.sum(q ** 2, 0), 1.0, atol=1e-12)
    np.testing.assert_array_equal(sh[0], 0.0)
    qa = util.augment(q)
    CG1, CG, nG, S2, S20_th, S20, NC = ot.op(qa, 0.12, 20, False, 50)
    assert nG == int(np.floor(4 * np.pi / 0.12 ** 2))
    assert sum(len(c) for c in CG1) == 2 * 14491
    assert all(len(c) <= 50 for c in CG) and len(CG) == S20_th.shape[1]
