"""zyz rotation of the incident frame (InitRotation, src/crosssec.c:82-118), test-side copy for the oracle tests."""
import numpy as np


def rotate_zyz(alpha, beta, gamma, v):
    a, b, g = np.radians([alpha, beta, gamma])
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    m = np.array([[ca * cb * cg - sa * sg, sa * cb * cg + ca * sg, -sb * cg],
                  [-ca * cb * sg - sa * cg, -sa * cb * sg + ca * cg, sb * sg],
                  [ca * sb, sa * sb, cb]])
    return m @ np.asarray(v, np.float64)
