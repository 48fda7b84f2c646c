#!/usr/bin/env python3
"""Build-time check of the device assembly for misplaced VGPR spills.

A divergent branch or loop on AMDGPU ends in a join block whose first instruction restores the wave's
execution mask (s_or_b64 exec, exec, sN).  A VGPR spill (scratch_store) that the register allocator
places in such a block *before* that restore runs with the narrowed mask -- with an empty one on the
exit edge of a divergent loop -- so the lanes outside it never reach memory, and a later reload under
the full mask returns stale data for them.  The fault is silent (no trap, plausible values).

The kernels of this library are built with the machine-level loop-invariant code motion pass disabled
(-mllvm -disable-machine-licm), which keeps the allocator from parking exp()/division constants in
long-lived VGPRs and removes these spills; this script guards that property: it exits with status 1,
naming the kernels and lines, when a scratch_store precedes the exec restore of its block with no
exec write, label or branch in between.  The check is conservative: a spill inside an if-body that
shares a block with the join is flagged too, although it is correct when the spilled value is
defined inside that body.

    python3 check_spill_placement.py <device assembly .s>
"""
import re
import sys


def find(lines):
    func, pending, hits = None, [], []
    for i, line in enumerate(lines):
        m = re.match(r'^(_Z\S+):', line)
        if m:
            func, pending = m.group(1), []
            continue
        if re.match(r'^(\.LBB\d+_\d+:|; %bb\.\d+:)', line) or re.match(r'^\s+s_(cbranch|branch|endpgm)', line):
            pending = []
            continue
        if re.match(r'^\s+s_or_b64 exec, exec, ', line):
            hits += [(func, j, i + 1) for j in pending]
            pending = []
            continue
        if re.match(r'^\s+s_\w+\s+(exec\b|exec,)|^\s+s_\w*saveexec', line):
            pending = []
            continue
        if re.match(r'^\s+scratch_store', line):
            pending.append(i + 1)
    return hits


def main(path):
    with open(path) as f:
        hits = find(f.read().split('\n'))
    for func, store, restore in hits[:20]:
        print(f"{path}:{store}: spill before the exec restore at line {restore} in {func}", file=sys.stderr)
    if hits:
        print(f"{len(hits)} misplaced VGPR spill(s): the device code would lose lanes of spilled values",
              file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
