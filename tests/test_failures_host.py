"""CPU checks of the failure-reporting host logic: every THR_ERR_* status of
the C ABI reaches Python as an exception naming the call, with the solver
statuses as their own classes (no compute call, no GPU needed)."""
import re
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_status_codes_match_header():
    import thirring_hip as G
    text = open(os.path.join(ROOT, "include", "thirring_hip.h")).read()
    codes = {m.group(1).lower(): int(m.group(2))
             for m in re.finditer(r"#define THR_(?:ERR_)?(OK|[A-Z]+)\s+(\d+)\b", text)
             if m.group(1) in ("OK", "ARG", "HIP", "NOCONV", "SIZE", "STATE", "NONFINITE")}
    assert codes == G.STATUS


@pytest.mark.parametrize("status,cls", [("noconv", "ThrNoConvergence"),
                                        ("nonfinite", "ThrNonFinite"), ("arg", "ThrError"),
                                        ("size", "ThrError")])
def test_status_raises_its_class(status, cls):
    import thirring_hip as G
    with pytest.raises(getattr(G, cls)) as e:
        G._check(G.STATUS[status], "some_call")
    assert e.value.status == G.STATUS[status]
    assert "some_call" in str(e.value)
    assert isinstance(e.value, G.ThrError)


def test_ok_status_is_silent():
    import thirring_hip as G
    G._check(0, "some_call")
