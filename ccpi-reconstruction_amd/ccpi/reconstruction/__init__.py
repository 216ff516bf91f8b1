"""``ccpi.reconstruction``: parallel- and cone-beam projectors and the
CGLS / SIRT / MLEM iteration bodies on one MI355X (gfx950).

Module paths mirror the reference's Python package
(Wrappers/Python/ccpi/reconstruction, modules ``parallelbeam`` and
``conebeam`` built from Wrappers/Python/src/*_module.cpp)."""
