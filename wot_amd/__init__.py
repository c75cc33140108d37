"""wot_amd: wot's unbalanced entropic optimal-transport path on AMD MI355X.

The ``wot_amd.ot`` namespace mirrors ``wot.ot`` (OTModel, transport_stablev2,
optimal_transport_duality_gap, compute_transport_matrix) and runs every solve
in hand-written HIP kernels (libwotk.so, C ABI in include/wotk.h).
"""
__version__ = '0.1.0'
