"""Windows of the 3D, the cylindrical and the 2D driver with a dielectric
for the reverse-mode tests: (Config keywords, steps). The 3D one is an 8 mm
cube of 2^3 base boxes (408 boxes); the cylindrical one 16 x 32 mm with the
seed on the axis (46 boxes) at 4 MV/m, so that the electron count depends
on the parameters well above its rounding (at 2 MV/m its directional
derivatives are 2.5e-5 of it); the dielectric one the reference's
dielectric (eps 1.5 below y = 8 mm) with the seed at (16, 12) mm, 120 boxes
of which the field solve tags those with eps."""

WINDOWS = {
    "3d": (dict(variant="3d", domain_len=8e-3, n_base=2, lvls_max=4,
                seed_r0=[[4e-3, 4e-3, 5e-3]], seed_r1=[[4e-3, 4e-3, 6e-3]],
                seed_width=[0.6e-3], applied_fld=2.5e6, n_boxes=4000, adjoint=True), 4),
    "cyl": (dict(variant="cyl", domain_len=16e-3, n_base=(1, 2), lvls_max=8,
                 seed_r0=[[0.0, 24e-3]], seed_r1=[[0.0, 24e-3]], applied_fld=4.0e6,
                 adjoint=True), 6),
    "2d_dielectric": (dict(variant="2d", lvls_max=6, epsilon_diel=1.5,
                           seed_r0=[[16e-3, 12e-3]], seed_r1=[[16e-3, 12e-3]], adjoint=True), 4),
}
