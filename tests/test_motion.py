"""CPU checks of the host-side kinematics (Python dual numbers and the native tvlm_motion_pitch_heave)
against the oracle's dual-number chain and against the closed form of tests/ut_kinematics.cpp."""
import math

import numpy as np

F32 = np.float32


def test_native_motion_matches_oracle(oracle):
    from tinyvlm_amd.uvlm_plan import motion_pitch_heave

    u, h, hp, a, ap, om, xp = 1.0, 0.1, 0.0, 0.0349066, math.pi / 2, 1.0, 0.25
    ts, dt = 30, 0.1
    T, dT, fs, ref = motion_pitch_heave(2, ts, dt, u, h, hp, a, ap, om, xp)
    mot = [u, h, hp, a, ap, om, xp]
    for k in range(ts):
        tk = 0.0 if k == 0 else (k - 1) * dt + dt
        To, _ = oracle.motion_transform(mot, tk)
        _, dTo = oracle.motion_transform(mot, k * dt)
        for m in range(2):
            assert np.allclose(T[k, m].reshape(4, 4, order="F"), To, rtol=0, atol=1e-15)
            assert np.allclose(dT[k, m].reshape(4, 4, order="F"), dTo, rtol=0, atol=1e-15)
        Tk, _ = oracle.motion_transform(mot, k * dt)
        assert np.allclose(ref[k], (Tk @ [0.5, 0, 0, 1])[:3], atol=1e-15)
        assert np.array_equal(fs[k], [u, 0, 0])


def test_python_kinematics_matches_oracle(oracle):
    from tinyvlm_amd.kinematics import pitch_heave

    mot = [1.3, 0.1, 0.2, 0.05, 0.4, 0.9, 0.25]
    _, fs, pt = pitch_heave(*mot)
    for t in (0.0, 0.37, 5.1):
        To, dTo = oracle.motion_transform(mot, t)
        D = pt.transform_dual(t)
        assert np.allclose(D.value(), To, atol=1e-15)
        assert np.allclose(D.grad(), dTo, atol=1e-15)
        assert np.allclose(fs.linear_velocity(t, (0, 0, 0)), [-1.3, 0, 0])


def test_kinematics_angular_velocity_kat():
    # tests/ut_kinematics.cpp:13-37: pitch about the origin by 0.5 pi sin(w t) after a freestream
    # translation; angular velocity y = 0.5 pi w cos(w t) (tolerance 1e-6 there)
    from tinyvlm_amd.kinematics import KinematicsTree, dsin, rotation, translation

    u, om = 1.0, 0.5
    tree = KinematicsTree()
    init = tree.add(lambda t: translation(0.0, 10.0, 0.0))
    fs = tree.add(lambda t: translation(-u * t, 0.0, 0.0)).after(init)
    pitch = tree.add(lambda t: rotation((0, 0, 0), (0, 1, 0), 0.5 * float(F32(3.141593)) * dsin(om * t))).after(fs)
    t = 1.0
    D = pitch.transform_dual(t)
    R, Rd = D.value()[:3, :3], D.grad()[:3, :3]
    W = Rd @ R.T
    ang = np.array([W[2, 1], W[0, 2], W[1, 0]])
    assert abs(ang[0]) < 1e-6 and abs(ang[2]) < 1e-6
    assert abs(ang[1] - 0.5 * float(F32(3.141593)) * om * math.cos(om * t)) < 1e-6
    v = pitch.linear_velocity(t, (1.0, 0.0, 0.0))
    th = 0.5 * float(F32(3.141593)) * math.sin(om * t)
    thd = 0.5 * float(F32(3.141593)) * om * math.cos(om * t)
    assert np.allclose(v, [-u - math.sin(th) * thd, 0.0, -math.cos(th) * thd], atol=1e-12)


def test_hb_poses_over_an_array_of_times_equal_one_time_at_a_time():
    """HBVLM._poses evaluates the HB kinematics tree (free stream, heave, pitch, flap) over all the
    collocation times in one pass of the dual-number arithmetic on numpy arrays; each time's
    transform and dT/dt equal the scalar evaluation of KinematicNode.transform_dual at that time
    (elementwise the same operations; 1e-15 allows a vectorised sin/cos that differs in the last
    bit)."""
    from tinyvlm_amd.hb import HBVLM, Vars, build_kinematics

    H = 5
    X = np.zeros((3, 2 * H + 1))
    X[0, 1], X[1, 2], X[2, 3], X[0, 4] = 0.1, math.radians(1.0), 0.02, 0.01
    fs, nodes = build_kinematics(X, 0.8, Vars(6.0), H)
    times = [0.7 + 0.013 * i for i in range(2 * H + 1)]
    poses = HBVLM._poses(nodes + [fs], times)
    for i, t in enumerate(times):
        for m, n in enumerate(nodes + [fs]):
            d = n.transform_dual(t)
            assert np.max(np.abs(poses[i][m][0] - d.value())) <= 1e-15
            assert np.max(np.abs(poses[i][m][1] - d.grad())) <= 1e-15
