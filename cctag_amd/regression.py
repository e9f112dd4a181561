"""The reference's regression log and its comparison rules.

A FileLog holds, for one input file, the detection parameters and per frame
the identified tags (id, status, x, y, quality) with the time the detection
took (src/applications/regression/TestLog.h:17-96); it is stored as a
boost::archive XML document with the top-level element "FileLog"
(TestLog.cpp:40-54).  compare() applies the checker's rules
(Regression.cpp:114-170, 221-235): same file name and ring count, same frame
count, frames in increasing order, per frame the same number of tags, then
only identified tags (status 1), sorted by id, with no duplicate ID, the same
status, the same ID and positions within `epsilon` px on each axis.
"""
from __future__ import annotations

import dataclasses
import xml.etree.ElementTree as ET

from . import _lib


class CheckError(Exception):
    """A regression check failed (the reference's check_error)."""


@dataclasses.dataclass
class DetectedTag:
    id: int
    status: int
    x: float
    y: float
    quality: float

    @staticmethod
    def of(marker: dict) -> "DetectedTag":
        return DetectedTag(int(marker["id"]), int(marker["status"]), float(marker["x"]),
                           float(marker["y"]), float(marker["quality"]))


@dataclasses.dataclass
class FrameLog:
    frame: int
    elapsed_time: float
    tags: list

    @staticmethod
    def detect(frame: int, img, detector) -> "FrameLog":
        """FrameLog::detect (TestLog.cpp:23-35) with a Detector."""
        import time
        t0 = time.perf_counter()
        markers = detector.detect(img, frame=frame)
        dt = time.perf_counter() - t0
        return FrameLog(frame, dt, [DetectedTag.of(m) for m in markers])


@dataclasses.dataclass
class FileLog:
    filename: str
    parameters: "_lib.Params"
    frame_logs: list

    def to_xml(self) -> str:
        pxml = _lib.params_to_xml(self.parameters)
        body = pxml[pxml.index(">", pxml.index("<CCTagsParams")) + 1:pxml.index("</CCTagsParams>")]
        fields = "".join("\t" + ln.strip() + "\n" for ln in body.splitlines() if ln.strip())
        out = ['<?xml version="1.0" encoding="UTF-8" standalone="yes" ?>',
               "<!DOCTYPE boost_serialization>",
               '<boost_serialization signature="serialization::archive" version="17">',
               '<FileLog class_id="0" tracking_level="0" version="0">',
               "\t<filename>%s</filename>" % _escape(self.filename),
               '\t<parameters class_id="1" tracking_level="0" version="0">']
        out += ["\t" + ln for ln in fields.rstrip("\n").split("\n")]
        out += ["\t</parameters>", "\t<frameLogs>",
                "\t\t<count>%d</count>" % len(self.frame_logs), "\t\t<item_version>0</item_version>"]
        for i, fl in enumerate(self.frame_logs):
            attrs = ' class_id="2" tracking_level="0" version="0"' if i == 0 else ""
            out += ["\t\t<item%s>" % attrs, "\t\t\t<frame>%d</frame>" % fl.frame,
                    "\t\t\t<elapsedTime>%.9g</elapsedTime>" % fl.elapsed_time, "\t\t\t<tags>",
                    "\t\t\t\t<count>%d</count>" % len(fl.tags),
                    "\t\t\t\t<item_version>0</item_version>"]
            for j, t in enumerate(fl.tags):
                a = ' class_id="3" tracking_level="0" version="0"' if (i, j) == (0, 0) else ""
                out += ["\t\t\t\t<item%s>" % a, "\t\t\t\t\t<id>%d</id>" % t.id,
                        "\t\t\t\t\t<status>%d</status>" % t.status,
                        "\t\t\t\t\t<x>%.9g</x>" % t.x, "\t\t\t\t\t<y>%.9g</y>" % t.y,
                        "\t\t\t\t\t<quality>%.9g</quality>" % t.quality, "\t\t\t\t</item>"]
            out += ["\t\t\t</tags>", "\t\t</item>"]
        out += ["\t</frameLogs>", "</FileLog>", "</boost_serialization>", "", ""]
        return "\n".join(out)

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_xml())

    @staticmethod
    def from_xml(text: str) -> "FileLog":
        root = ET.fromstring(text.split("<!DOCTYPE boost_serialization>")[-1].strip())
        fl = root.find("FileLog")
        if fl is None:
            raise CheckError("no FileLog element")
        params_el = fl.find("parameters")
        inner = "".join(ET.tostring(c, encoding="unicode") for c in params_el)
        params = _lib.params_from_xml("<CCTagsParams>%s</CCTagsParams>" % inner)
        frames = []
        for item in fl.find("frameLogs").findall("item"):
            tags = [DetectedTag(int(t.findtext("id")), int(t.findtext("status")),
                                float(t.findtext("x")), float(t.findtext("y")),
                                float(t.findtext("quality")))
                    for t in item.find("tags").findall("item")]
            frames.append(FrameLog(int(item.findtext("frame")), float(item.findtext("elapsedTime")),
                                   tags))
        return FileLog(fl.findtext("filename") or "", params, frames)

    @staticmethod
    def load(path: str) -> "FileLog":
        with open(path) as f:
            return FileLog.from_xml(f.read())


def _escape(s: str) -> str:
    return s.replace("&", "&amp;").replace("<", "&lt;").replace(">", "&gt;")


def _sort_tags(tags):
    """SortTags (Regression.cpp:221-235): identified tags by id; False on a
    duplicate ID."""
    kept = sorted((t for t in tags if t.status >= 1), key=lambda t: t.id)
    return kept, all(a.id != b.id for a, b in zip(kept, kept[1:]))


def compare(reference: FileLog, test: FileLog, epsilon: float = 0.5) -> dict:
    """TestChecker::compare: raises CheckError with the reference's message
    on the first failed rule; returns the elapsed-time and quality
    differences (test - reference) it accumulates."""
    if reference.filename != test.filename:
        raise CheckError("mismatching filenames")
    if reference.parameters.n_crowns != test.parameters.n_crowns:
        raise CheckError("mismatching parameters")
    if len(reference.frame_logs) != len(test.frame_logs):
        raise CheckError("mismatching frame counts")
    for name, log in (("reference", reference), ("test", test)):
        if any(a.frame > b.frame for a, b in zip(log.frame_logs, log.frame_logs[1:])):
            raise CheckError("%s log frames not monotonic" % name)
    elapsed, quality = [], []
    for i, (r, t) in enumerate(zip(reference.frame_logs, test.frame_logs)):
        if len(r.tags) != len(t.tags):
            raise CheckError("different # of tags in frame %d" % i)
        rt, ok_r = _sort_tags(r.tags)
        tt, ok_t = _sort_tags(t.tags)
        if not ok_r:
            raise CheckError("reference log contains duplicate IDs")
        if not ok_t:
            raise CheckError("test log contains duplicate IDs")
        elapsed.append(t.elapsed_time - r.elapsed_time)
        if len(rt) != len(tt):
            raise CheckError("tags of different status in frame %d" % i)
        for a, b in zip(rt, tt):
            quality.append(b.quality - a.quality)
            if a.id != b.id:
                raise CheckError("tags with different IDs frame %d" % i)
            if abs(a.x - b.x) > epsilon or abs(a.y - b.y) > epsilon:
                raise CheckError("tags at different positions in frame %d" % i)
    return {"elapsed_diff": elapsed, "quality_diff": quality}
