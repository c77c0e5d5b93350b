"""Shared tolerance of the multi-step trajectory comparisons.

The implicit-Euler step ends its Newton loop once the RMS Newton step of the last three
iterations is at most PNTol * dt (IMPLICIT_EULER.h:647-710, 866).  Two correct runs whose
floating-point rounding differs (a different summation order in a reduction, FP32 matrix
blocks) can take one Newton iteration more or less, or double the barrier stiffness one
iteration apart, so their end-of-step states agree only to about that stopping resolution.
The bar is therefore the larger of 1e-6 of the step's displacement and PNTol * dt."""


def trajectory_bar(sd, disp):
    return max(1e-6 * max(disp, 1e-3), sd.params["PNTol"] * sd.params["dt"])
