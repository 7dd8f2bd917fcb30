"""Closed-form whole-space solutions used to pin the oracle's source injection.

Homogeneous isotropic full space, point sources with a Ricker time function,
Aki & Richards, Quantitative Seismology (2nd ed.):

* moment-tensor source, eq. (4.29): displacement u_n for a moment history
  M_pq(t);
* single force, eq. (4.23): displacement u_i for a force history X_0(t) in
  direction j.

The velocity-stress scheme injects the moment *rate* into the stress rate
(dT_ij/dt -= Mdot_ij(t) delta(x)) and a force into the velocity rate
(dV_i/dt += F_i(t) delta(x) / rho), so with a Ricker history R(t) in the
source tables the recorded velocities are the time derivatives of the two
formulas with M(t) = M int R and X_0(t) = F R:

    moment: v_n = M_pq [ N/(4 pi rho r^4) I1(t) + IP/(4 pi rho a^2 r^2) R(t - r/a)
                         - IS/(4 pi rho b^2 r^2) R(t - r/b)
                         + FP/(4 pi rho a^3 r) R'(t - r/a) - FS/(4 pi rho b^3 r) R'(t - r/b) ]
    force:  v_i = F_j [ (3 g_i g_j - d_ij)/(4 pi rho r^3) I2(t)
                        + g_i g_j/(4 pi rho a^2 r) R'(t - r/a)
                        - (g_i g_j - d_ij)/(4 pi rho b^2 r) R'(t - r/b) ]

with I1 = int_{r/a}^{r/b} tau R(t - tau) dtau and I2 = the same integral of
R'. Every antiderivative of the Ricker wavelet is elementary, so nothing
here is integrated numerically.
"""
import math

import numpy as np


class Ricker:
    """R(t) = (1 - 2 s^2) exp(-s^2), s = pi f0 (t - tc)  (= (1 - u^2/2) exp(-u^2/4), u = 2 pi f0 (t - tc))."""

    def __init__(self, f0, tc):
        self.f0, self.tc = float(f0), float(tc)
        self.p = math.pi * self.f0

    def R(self, t):
        s = self.p * (t - self.tc)
        return (1 - 2 * s * s) * np.exp(-s * s)

    def dR(self, t):
        s = self.p * (t - self.tc)
        return 2 * self.p * s * (2 * s * s - 3) * np.exp(-s * s)

    def Q(self, t):
        """Antiderivative of R vanishing at -infinity: (t - tc) exp(-s^2)."""
        s = self.p * (t - self.tc)
        return (t - self.tc) * np.exp(-s * s)

    def sQ(self, t):
        """Antiderivative of t R(t): t Q(t) - int Q = t Q(t) + exp(-s^2) / (2 p^2)."""
        s = self.p * (t - self.tc)
        return t * self.Q(t) + np.exp(-s * s) / (2 * self.p * self.p)

    def near_R(self, t, ta, tb):
        """int_{ta}^{tb} tau R(t - tau) dtau = int_{t-tb}^{t-ta} (t - s) R(s) ds."""
        s1, s2 = t - tb, t - ta
        return t * (self.Q(s2) - self.Q(s1)) - (self.sQ(s2) - self.sQ(s1))

    def near_dR(self, t, ta, tb):
        """int_{ta}^{tb} tau R'(t - tau) dtau = [(t - s) R(s)]_{s1}^{s2} + int_{s1}^{s2} R."""
        s1, s2 = t - tb, t - ta
        return (t - s2) * self.R(s2) - (t - s1) * self.R(s1) + self.Q(s2) - self.Q(s1)


def moment_velocity(M, src, rcv, t, alpha, beta, rho, stf):
    """Velocity (3, nt) at rcv from a moment-rate source M (3x3) * stf.R at src (Aki & Richards 4.29, d/dt)."""
    d = np.asarray(rcv, float) - np.asarray(src, float)
    r = float(np.linalg.norm(d))
    g = d / r
    I = np.eye(3)
    ggg = np.einsum("n,p,q->npq", g, g, g)
    gn_dpq = np.einsum("n,pq->npq", g, I)
    gp_dnq = np.einsum("p,nq->npq", g, I)
    gq_dnp = np.einsum("q,np->npq", g, I)
    N = 15 * ggg - 3 * gn_dpq - 3 * gp_dnq - 3 * gq_dnp
    IP = 6 * ggg - gn_dpq - gp_dnq - gq_dnp
    IS = 6 * ggg - gn_dpq - gp_dnq - 2 * gq_dnp
    FP = ggg
    FS = np.einsum("np,q->npq", np.outer(g, g) - I, g)
    c = lambda T: np.einsum("npq,pq->n", T, M)  # noqa: E731
    ta, tb = r / alpha, r / beta
    k = 4 * math.pi * rho
    return (np.outer(c(N), stf.near_R(t, ta, tb)) / (k * r ** 4)
            + np.outer(c(IP), stf.R(t - ta)) / (k * alpha ** 2 * r ** 2)
            - np.outer(c(IS), stf.R(t - tb)) / (k * beta ** 2 * r ** 2)
            + np.outer(c(FP), stf.dR(t - ta)) / (k * alpha ** 3 * r)
            - np.outer(c(FS), stf.dR(t - tb)) / (k * beta ** 3 * r))


def force_velocity(F, src, rcv, t, alpha, beta, rho, stf):
    """Velocity (3, nt) at rcv from a force F (3,) * stf.R at src (Aki & Richards 4.23, d/dt)."""
    d = np.asarray(rcv, float) - np.asarray(src, float)
    r = float(np.linalg.norm(d))
    g = d / r
    I = np.eye(3)
    G = np.outer(g, g)
    ta, tb = r / alpha, r / beta
    k = 4 * math.pi * rho
    F = np.asarray(F, float)
    return (np.outer((3 * G - I) @ F, stf.near_dR(t, ta, tb)) / (k * r ** 3)
            + np.outer(G @ F, stf.dR(t - ta)) / (k * alpha ** 2 * r)
            - np.outer((G - I) @ F, stf.dR(t - tb)) / (k * beta ** 2 * r))
