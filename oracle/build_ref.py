"""Reference build for the oracle: not buildable in this image.

MOLGW's hot path is Fortran whose modules need the integral libraries
(LIBCINT/LIBINT, absent here) to link and MPI/ScaLAPACK-free builds of
generated sources; building it would require stand-ins for missing
libraries, which is not allowed. The oracle is therefore pinned by the
reference's golden values and by closed forms (tests/test_oracle.py), and
this build step does nothing.
"""


def build() -> None:
    return None
