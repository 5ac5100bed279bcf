"""Mesh cases shared by the golden-fixture writer and the parity tests."""
from elastic_amd import mesh as M

# name -> (mesh builder, solver tolerance, harness mode, fixture keeps full arrays)
# (harness mode 47 = PCG, GMRES, BiCGSTAB, 3x3 block Jacobi, coloured
# assembly, and the nodal stresses with the result files; + 256 on the 2-D
# meshes = 2x2 block Jacobi)
CASES = {
    "hex8_4": (lambda: M.hex8_cube(4), 1e-11, 47, True),
    "hex8_5j": (lambda: M.hex8_cube(5, jitter=0.15, seed=M.BASE_SEED + 2), 1e-10, 47, True),
    "tet4_3j": (lambda: M.tet4_cube(3, jitter=0.2, seed=M.BASE_SEED + 3), 1e-10, 47, True),
    "quad4_8": (lambda: M.quad4_square(8, E=1.3, nu=0.27, traction=(0.3, -0.7)), 1e-11, 303, True),
    "tria3_8": (lambda: M.tria3_square(8, E=0.8, nu=0.33, traction=(-0.2, 0.9)), 1e-11, 303, True),
    "quad4_6ps": (lambda: M.quad4_square(6, nu=0.25, plane_stress=True, thickness=0.5, edge="top"),
                  1e-11, 303, True),
    "tria3_6ps": (lambda: M.tria3_square(6, nu=0.25, plane_stress=True, thickness=0.5, edge="top"),
                  1e-11, 303, True),
    "req0": (lambda: M.serving_request(0, nmin=8, nmax=16), 1e-10, 257, True),
    "req1": (lambda: M.serving_request(1, nmin=8, nmax=16), 1e-10, 257, True),
    # workload 1 (hex8 20^3) at its full size: digests + solution vector
    "hex8_20": (lambda: M.config_mesh(1), 1e-10, 13, False),
    # the reference's own quad deck (input/quad/exemplo1_apostila_el.dat) is
    # this generated mesh node for node; its fixture is written from the deck
    "quad_el": (lambda: M.quad4_square(100, E=1.0, nu=0.3, traction=(0.0, -1.0), edge="top"),
                1e-11, 1, False),
}

# full-size workloads whose reference set-up (RCM, numeq, CSRC, colouring,
# coloured assembly, one product) is kept as digests: hex8 100^3 jittered,
# tet4 80^3 x 6 and the plasticity mesh (hex8 60^3 between grips)
DIGESTS = {
    "cfg2_hex8_100": lambda: M.config_mesh(2),
    "cfg3_tet4_80": lambda: M.config_mesh(3),
    "cfg5_pull_60": lambda: M.hex8_block_pull(60, 5e-3)[0],
}

# fixtures written from the reference's own decks
REF_DECKS = {
    "quad_el": "quad/exemplo1_apostila_el.dat",
    "quad_nl": "quad/exemplo1_apostila_nl.dat",
    "quad_nf": "quad/exemplo1_apostila_nf.dat",
    "quad_el_ept": "quad/exemplo1_apostila_el_ept.dat",
}

SMALL = [k for k, v in CASES.items() if v[3]]
