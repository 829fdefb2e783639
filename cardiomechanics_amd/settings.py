"""Reference-style XML settings files as dotted keys (ParameterMap,
mechanics/src/ParameterMap/ParameterMap.h:29-152): <Solver><Type>NewmarkBeta</Type>
becomes {"Solver.Type": "NewmarkBeta"}; the root tag (<settings>) is dropped and the
values stay text.  The same reading as the host C++ layer's ParameterMap::ParseXML.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET


def parse_xml_string(text: str) -> dict:
    root = ET.fromstring(text)
    out = {}

    def walk(node, prefix):
        for ch in node:
            key = f"{prefix}.{ch.tag}" if prefix else ch.tag
            if len(ch):
                walk(ch, key)
            else:
                out[key] = (ch.text or "").strip()

    walk(root, "")
    return out


def parse_xml(path: str) -> dict:
    with open(path) as fh:
        return parse_xml_string(fh.read())


def as_bool(v) -> bool:
    """ParameterMap's bool conversion: true / 1 / True / TRUE"""
    if isinstance(v, str):
        return v.strip() in ("true", "1", "True", "TRUE")
    return bool(v)
