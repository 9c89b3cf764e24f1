"""VCtl from a PSD's attribute files, as the scan set-up derives it before
the sequence is compiled (Src/Main/DoPreScan.m:33-122).

The reference reads its simulation panel (the PSD folder's SimuAttr.xml and
the technique's attribute file: EPI.xml, FSE.xml, Radial.xml, Spiral.xml,
...) into VCtl and then fixes the encoding geometry per technique: the
readout gets an odd sample count so that one sample sits on kx = 0, the
phase and slice counts are made even, RFreq/RPhase/FOVSlice follow, and
FirstPhNum/SecondPhNum name the TR loop of the sequence compiler. This
module does the same on plain dicts; the result goes to
mrilab_amd.engine.compile_psd.
"""
import math
import re
import xml.etree.ElementTree as ET


def read_attrs(xml_text: str) -> dict:
    """Attributes of every element below the root of an attribute file.
    A "$k'a','b'" option list takes its k-th entry (the panel's popup); a
    number is parsed; anything else stays text."""
    out = {}
    root = ET.fromstring(xml_text)
    for el in [root] + list(root.iter()):
        for k, v in el.attrib.items():
            if el is root and k in ("Name", "Notes", "Description"):
                continue
            out[k] = _value(v)
    return out


def _value(v: str):
    if v.startswith("$"):
        opts = re.findall(r"'([^']*)'", v[2:])
        return opts[int(v[1]) - 1]
    try:
        return float(v)
    except ValueError:
        pass
    # a list of numbers, as the panel's str2num reads "5e-3 5e-3" (ME_TEs)
    parts = v.replace(",", " ").split()
    try:
        return [float(x) for x in parts] if len(parts) > 1 else v
    except ValueError:
        return v


def derive(attrs: dict) -> dict:
    """The technique-specific VCtl fields (DoPreScan.m:33-109) and
    TEAnchorTime = 0 (DoPreScan.m:111)."""
    v = dict(attrs)
    if "R_SpokeNum" in v or "S_ShotNum" in v:  # radial, spiral: square pixels
        v["RFreq"] = v["FOVFreq"] / v["ResFreq"]
        v["FOVPhase"] = v["FOVFreq"]
        v["ResPhase"] = v["ResFreq"]
        v["RPhase"] = v["FOVPhase"] / v["ResPhase"]
        first = v["R_SpokeNum"] if "R_SpokeNum" in v else v["S_ShotNum"]
    else:  # EPI, FSE, Cartesian: odd readout count, even phase count
        v["ResFreq"] = v["ResFreq"] + (1 if math.fmod(v["ResFreq"], 2) == 0 else 0)
        v["RFreq"] = v["FOVFreq"] / (v["ResFreq"] - 1)
        if "EPI_ShotNum" in v:
            v["EPI_ETL"] = v["EPI_ETL"] - math.fmod(v["EPI_ETL"], 2)
            v["ResPhase"] = v["EPI_ETL"] * v["EPI_ShotNum"]
            first = v["EPI_ShotNum"]
        elif "FSE_ShotNum" in v:
            v["FSE_ETL"] = v["FSE_ETL"] - math.fmod(v["FSE_ETL"], 2)
            v["ResPhase"] = v["FSE_ETL"] * v["FSE_ShotNum"]
            first = v["FSE_ShotNum"]
        else:
            v["ResPhase"] = v["ResPhase"] - math.fmod(v["ResPhase"], 2)
            first = v["ResPhase"]
        v["RPhase"] = v["FOVPhase"] / v["ResPhase"]
    v["SliceNum"] = max(1.0, v["SliceNum"] - math.fmod(v["SliceNum"], 2))
    v["RSlice"] = v["SliceThick"]
    v["FOVSlice"] = v["SliceNum"] * v["SliceThick"]
    v["FirstPhNum"] = first
    v["SecondPhNum"] = v["SliceNum"]
    v["TEAnchorTime"] = 0.0
    return v


def vctl_from_files(*xml_texts: str, **overrides) -> dict:
    """VCtl from the attribute files of a PSD (later files win), with
    keyword overrides applied before the derivation."""
    attrs = {}
    for t in xml_texts:
        attrs.update(read_attrs(t))
    attrs.update(overrides)
    return derive(attrs)
