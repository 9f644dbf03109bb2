"""The product package never reaches the oracle: no module under summit_amd/
imports, reads or runs anything from oracle/ (the oracle is the checker only)."""
from __future__ import annotations

import ast
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_product_modules_do_not_import_the_oracle():
    pkg = os.path.join(ROOT, "summit_amd")
    offenders = []
    for dirpath, _, files in os.walk(pkg):
        for name in files:
            if not name.endswith(".py"):
                continue
            path = os.path.join(dirpath, name)
            tree = ast.parse(open(path, encoding="utf-8").read(), filename=path)
            for node in ast.walk(tree):
                mods = []
                if isinstance(node, ast.Import):
                    mods = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom) and node.module:
                    mods = [node.module]
                if any(m == "oracle" or m.startswith("oracle.") for m in mods):
                    offenders.append(os.path.relpath(path, ROOT))
            if "oracle/" in open(path, encoding="utf-8").read().replace("\\\\", "/"):
                offenders.append(os.path.relpath(path, ROOT) + " (path)")
    assert not offenders, offenders
