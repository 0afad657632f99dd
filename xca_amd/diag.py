"""Connected pairings (backbone topologies) and their parity signs: python/triqs_xca/diag.py:34-127.

Order 1: [(0, 1)] (+); order 2: [(0, 2), (1, 3)] (-); order 3: four topologies; order 4: 27."""


def _pairings(parity, partial, unpaired):
    if not unpaired:
        yield parity, partial
        return
    for i in range(1, len(unpaired)):
        pair = (unpaired[0], unpaired[i])
        rest = unpaired[1:i] + unpaired[i + 1:]
        yield from _pairings(-parity if i % 2 == 0 else parity, partial + [pair], rest)


def all_pairings(order):
    yield from _pairings(1, [], list(range(2 * order)))


def _crossing(p1, p2):
    return (p1[0] < p2[0] < p1[1] < p2[1]) or (p2[0] < p1[0] < p2[1] < p1[1])


def is_connected(pairing):
    """Connected in the sense of diag.py:110-127 with ks = [1]: starting from the lines that cross
    time 0+ (contain vertex 0), every line must be reachable through crossings."""
    stack = [p for p in pairing if (p[0] <= 0) != (p[1] <= 0)]
    rest = [p for p in pairing if (p[0] <= 0) == (p[1] <= 0)]
    while stack:
        pc = stack.pop(0)
        for idx in reversed(range(len(rest))):
            if _crossing(pc, rest[idx]):
                stack.append(rest.pop(idx))
    return not rest


def all_connected_pairings(order):
    for parity, pairing in all_pairings(order):
        if is_connected(pairing):
            yield parity, pairing
