"""Writes cornell_digest.json: the Cornell box vertex and face tables of the reference
(src/cornell_box.rs) as float32 / uint32 arrays, stored as their SHA-256 digests plus the
array shapes.  Run in the build container, where the reference checkout is readable."""
import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/cornell_box.rs")
text = SRC.read_text()
vblock = text[text.index("VERTICES"):text.index("INDICES")]
fblock = text[text.index("INDICES"):]
num = re.compile(r"-?\d+\.\d+|-?\d+")
vblock = vblock[vblock.index("= [") + 3:vblock.index("];")]
fblock = fblock[fblock.index("= [") + 3:fblock.index("];")]
strip = lambda b: "\n".join(l.split("//")[0] for l in b.splitlines())
verts = np.array([float(x) for x in num.findall(strip(vblock))], np.float32).reshape(-1, 3)
faces = np.array([int(x) for x in num.findall(strip(fblock))], np.uint32).reshape(-1, 3)
out = {"vertices_shape": list(verts.shape), "faces_shape": list(faces.shape),
       "vertices_sha256": hashlib.sha256(verts.tobytes()).hexdigest(),
       "faces_sha256": hashlib.sha256(faces.tobytes()).hexdigest()}
Path(__file__).with_name("cornell_digest.json").write_text(json.dumps(out, indent=1) + "\n")
print(out)
