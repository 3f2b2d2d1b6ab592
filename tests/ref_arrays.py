"""Reference-layout module arrays for tests (defined with the oracle)."""
from oracle.ref_tables import reference_arrays  # noqa: F401
