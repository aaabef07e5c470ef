"""Write the frame-output fixtures from the two frames the reference ships
as documentation (docs/running.md:33,55):

- docs/vortex-street.png, a vorticity map of the default run written by
  render::write_vorticity (src/render.rs:176-204) through png::write_rgb
  (src/png.rs:174-197). It is copied as is: its pixels (the zlib stream
  inflated) are the input, the file is the expected output.
- docs/velocity-arrows.svg, an arrow plot of a 2048 x 1280 run written by
  render::write_arrows (src/render.rs:209-267). It is reduced to its text
  lines outside the arrows (header, plate rectangles, closing tags) and the
  arrows' ten numbers each, in tenths, as int16.

Run from the repository root: python tests/golden/make_docs_frames.py
"""
import json
import os
import re
import shutil

import numpy as np

REF = "/root/reference/docs"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "docs_frames")

NUM = r"(-?\d+\.\d)"
PATH = re.compile(r'<path d="M{0} {0}L{0} {0}m{0} {0}L{0} {0}l{0} {0}"/>$'.format(NUM))


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "vortex-street.png"), os.path.join(OUT, "vortex-street.png"))
    lines = open(os.path.join(REF, "velocity-arrows.svg")).read().split("\n")
    other, paths = [], []
    for ln in lines:
        m = PATH.match(ln)
        if m:
            v = [round(float(x) * 10) for x in m.groups()]
            assert v[6:8] == v[2:4], ln  # the head is drawn back from the tip
            paths.append(v)
        elif ln.startswith("<path"):
            raise ValueError("unparsed arrow: " + ln)
        else:
            other.append(ln)
    arr = np.array(paths, dtype=np.int64)
    assert np.abs(arr).max() < 2**15
    np.save(os.path.join(OUT, "velocity-arrows-paths.npy"), arr.astype(np.int16))
    with open(os.path.join(OUT, "velocity-arrows.json"), "w") as f:
        json.dump({"source": "docs/velocity-arrows.svg", "lines_without_arrows": other,
                   "arrows": len(paths), "columns": "x0 y0 x1 y1 mx my x1 y1 lx ly, tenths"}, f,
                  indent=1)
    print(f"{len(paths)} arrows, {len(other)} other lines")


if __name__ == "__main__":
    main()
