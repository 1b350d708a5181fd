"""Writes feature_keys.json: the keys of the reference's FeatureMap and its
model codes, read as text from the reference's src/features.rs (the
fixture is data: key names and integers).

Literal keys come from `f.insert("...")` calls and the `for key in [...]`
lists; keys built with format! are expanded over the loops they sit in:
{key} / {offset} over 30d / 60d / 90d (src/features.rs:166-214),
{i} over 0..4 (:292-297) and {name} over model_name_str (:50-61, 327-334).

Run from the repository root: python tests/golden/make_feature_keys.py
"""
import json
import re
from pathlib import Path

SRC = Path("/root/reference/src/features.rs").read_text()
keys = set(re.findall(r'insert\(\s*"([a-z0-9_]+)"', SRC))
for block in re.findall(r"for key in \[(.*?)\]", SRC, flags=re.S):
    keys.update(re.findall(r'"([a-z0-9_]+)"', block))
names = re.findall(r'SviModelName::(\w+) => "(\w+)"', SRC)
model_names = [n for _, n in names]
codes = {v: float(c) for v, c in re.findall(r"SviModelName::(\w+) => ([0-9.]+),", SRC)}
for tmpl in re.findall(r'format!\("([a-z0-9_{}]+)"\)', SRC):
    if "{key}" in tmpl or "{offset}" in tmpl:
        keys.update(tmpl.replace("{key}", o).replace("{offset}", o) for o in ("30d", "60d", "90d"))
    elif "{i}" in tmpl:
        keys.update(tmpl.replace("{i}", str(i)) for i in range(4))
    elif "{name}" in tmpl:
        keys.update(tmpl.replace("{name}", n) for n in model_names)
    else:
        raise SystemExit(f"unexpanded template {tmpl}")
out = {"keys": sorted(keys), "model_codes": codes, "model_names": model_names}
Path(__file__).with_name("feature_keys.json").write_text(json.dumps(out, indent=1) + "\n")
print(len(keys), "keys,", len(codes), "model codes")
