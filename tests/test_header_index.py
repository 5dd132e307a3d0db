"""Every function include/mpcf_hip.h declares appears in the entry-point index
of the integration notes, by its full name or, inside a row that names its
family's first function, by its suffix (`_stats` for mpcf_server_stats)."""
import os
import re

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "mpcf_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpcf_[a-z0-9_]+)\s*\(", src)))


def index_rows():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    start = text.index("## 9. Entry-point index")
    return [l for l in text[start:].splitlines() if l.startswith("| `mpcf_")]


def test_every_declared_function_is_indexed():
    names = declared()
    assert len(names) > 80
    rows = index_rows()
    missing = []
    for name in names:
        if any(f"`{name}`" in r for r in rows):
            continue
        fam = next((p for p in ("mpcf_batch", "mpcf_server") if name.startswith(p + "_")), None)
        if fam and any(f"`{fam}_" in r and f"`{name[len(fam):]}`" in r for r in rows):
            continue
        missing.append(name)
    assert not missing, missing
