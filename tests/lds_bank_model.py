"""Model of the LDS bank conflicts of the line transforms in LDS (spr_fft.hpp).

The model used to choose the LDS layouts of the column tiles, the row passes and
the on-chip iteration kernel. An 8-byte LDS access by a 64-lane wavefront is
served 16 lanes per cycle (32 banks of 4 bytes, so 16 slots of 8 bytes); lanes
of one group that touch different addresses in the same slot serialise. The
model replays the element indices that `fft_line_p` stores and loads between its
stages (and, for the on-chip kernel, the loads and stores at the row/column phase
boundaries), maps them through a layout to LDS addresses, and counts the cycles
beyond the 4 a conflict-free access takes. Its counts agree with the
SQ_LDS_BANK_CONFLICT counter per wave (tests/test_lds_bank_model.py).

Layouts (addresses in float2 units; `line` is the column or row a lane works on):
  column tile, padded:  element y of column c at y * (C + pad) + c
  column tile, skewed:  element y of column c at (y + (y >> skew)) * C + c
  row:                  element x of row r at r * RS + x + (x >> shift)
"""
from __future__ import annotations

SLOTS = 16          # 8-byte bank slots
GROUP = 16          # lanes served per cycle for 8-byte accesses
WAVE = 64


def stage_accesses(n, v, radices, t):
    """Element indices thread t touches in the LDS exchanges of one transform,
    one list per LDS instruction, in program order: the store of stage s
    (s < last) and the load of stage s + 1 (spr_fft.hpp stage_store/load)."""
    T = n // v
    out = []
    span = 1
    for s, r in enumerate(radices):
        nb = v // r
        if s > 0:
            for u in range(nb):
                for k in range(r):
                    out.append(t + T * u + (n // r) * k)
        if s < len(radices) - 1:
            for u in range(nb):
                b = t + T * u
                base = (b // span) * span * r + b % span
                for q in range(r):
                    out.append(base + q * span)
        span *= r
    return out


def access_cycles(addresses):
    """Cycles of one 8-byte LDS instruction of a wavefront (64 addresses)."""
    cycles = 0
    for g in range(0, len(addresses), GROUP):
        slots = {}
        for a in addresses[g:g + GROUP]:
            slots.setdefault(a % SLOTS, set()).add(a)
        cycles += max(len(s) for s in slots.values())
    return cycles


def extra_cycles(instructions):
    """Conflict cycles of a list of instructions (each: 64 lane addresses)."""
    return sum(access_cycles(a) - len(a) // GROUP for a in instructions)


def _transform_instructions(n, v, radices, lanes, address):
    per_lane = [stage_accesses(n, v, radices, t) for _, t in lanes]
    return [[address(lanes[i][0], per_lane[i][k]) for i in range(WAVE)] for k in range(len(per_lane[0]))]


def column_pass(n, v, radices, columns, pad=1, skew=0, transforms=2):
    """Conflict cycles per wave of a column pass (lane = t * C + c, as k_col)."""
    C = columns
    if skew:
        address = lambda c, y: (y + (y >> skew)) * C + c
    else:
        address = lambda c, y: y * (C + pad) + c
    waves = C * (n // v) // WAVE
    total = 0
    for w in range(waves):
        lanes = [((w * WAVE + l) % C, (w * WAVE + l) // C) for l in range(WAVE)]
        total += transforms * extra_cycles(_transform_instructions(n, v, radices, lanes, address))
    return total / waves


def row_pass(n, v, radices, shift, transforms=2):
    """Conflict cycles per wave of a row pass (lane = row * T + t, as k_row)."""
    T = n // v
    rs = n + (n >> shift)
    address = lambda r, x: r * rs + x + (x >> shift)
    rows_per_wave = max(1, WAVE // T)
    lanes = [(l // T, l % T) if T < WAVE else (0, l) for l in range(WAVE)]
    assert len({r for r, _ in lanes}) == rows_per_wave
    return transforms * extra_cycles(_transform_instructions(n, v, radices, lanes, address))


def fused_iteration(n, rs, shift, column_t_major=True):
    """Conflict cycles per wave of one iteration of the on-chip kernel
    (spr_fused.hpp): a column phase and a row phase, each with two transforms
    and the loads and stores at the phase boundary. Rows: lane = row * T + t;
    columns: lane = t * (64 / T) + column when column_t_major, else as rows."""
    v = 16
    T = n // v
    lpw = WAVE // T
    radices = {128: (16, 8), 64: (16, 4)}[n]
    row_addr = lambda r, x: r * rs + x + (x >> shift)
    col_addr = lambda c, y: y * rs + c + (c >> shift)
    waves = n * T // WAVE
    total = 0
    for w in range(waves):
        rows = [((w * WAVE + l) // T, (w * WAVE + l) % T) for l in range(WAVE)]
        cols = [(w * lpw + l % lpw, l // lpw) for l in range(WAVE)] if column_t_major else rows
        for lanes, address in ((rows, row_addr), (cols, col_addr)):
            boundary = [[address(line, t + T * j) for line, t in lanes] for j in range(v)]
            total += 2 * extra_cycles(_transform_instructions(n, v, radices, lanes, address))
            total += 2 * extra_cycles(boundary)
    return total / waves
