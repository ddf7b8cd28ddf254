"""The round-end smoke check (``__graft_entry__.smoke``) as a GPU test: one
exact-size RenderStep against the float64 oracle, so a change to the step's
outputs that the smoke relies on fails here first."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_point():
    import __graft_entry__

    __graft_entry__.smoke()
