"""Deterministic synthetic workloads (counter-based, numpy).

Streams are u64(i) = splitmix64(seed + i); independent sub-streams of one
seed start at hashed counters (sub_seed).  Uniform 31-bit keys are
u64 >> 33, values are MurmurHash3_x86_32(key, seed 0), widths are
log-uniform.  The paper-sweep key bijection follows
FliX/paper_lsmu_sweep.cu:64-67.
"""

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
    return x ^ (x >> np.uint64(31))


def stream(seed, n, offset=0):
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return splitmix64(idx + np.uint64(seed))


def _rotl(x, r):
    return (x << np.uint32(r)) | (x >> np.uint32(32 - r))


def murmur3_u32(keys, seed=0):
    k = np.asarray(keys, dtype=np.uint32)
    with np.errstate(over="ignore"):
        k = k * np.uint32(0xCC9E2D51)
        k = _rotl(k, 15)
        k = k * np.uint32(0x1B873593)
        h = np.full(k.shape, seed, dtype=np.uint32) ^ k
        h = _rotl(h, 13)
        h = h * np.uint32(5) + np.uint32(0xE6546B64)
        h = h ^ np.uint32(4)
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x85EBCA6B)
        h = h ^ (h >> np.uint32(13))
        h = h * np.uint32(0xC2B2AE35)
        h = h ^ (h >> np.uint32(16))
    return h


def log_uniform(x, a, b):
    u = (np.asarray(x, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) \
        * 2.0 ** -53
    la, lb = np.log2(a), np.log2(b)
    v = np.floor(np.exp2(la + u * (lb - la)))
    return np.clip(v, a, b).astype(np.uint64)


def paper_key_for_index(index):
    i = np.asarray(index, dtype=np.uint32)
    with np.errstate(over="ignore"):
        return (i * np.uint32(747796405) + np.uint32(289133645)) & \
            np.uint32((1 << 30) - 1)


# ---------------------------------------------------------------- configs
SEEDS = {1: 0xC0F10001, 2: 0xC0F10002, 3: 0xC0F10003, 4: 0xC0F10004,
         5: 0xC0F10005}


def sub_seed(seed, tag):
    """Start counter of an independent sub-stream: a hashed 64-bit value, so
    sub-streams of one seed never overlap in practice."""
    return int(splitmix64(np.uint64((int(seed) << 16) ^ int(tag))))


def uniform31(seed, n, offset=0):
    return (stream(seed, n, offset) >> np.uint64(33)).astype(np.uint32)


class LsmWorkload:
    """Plain-LSM workload: insert/erase batches, then lookups and ranges.

    keys in [0, 2^31); values murmur3(key).  In every batch after the
    first, entry i is a tombstone with probability `erase_fraction` (decided
    by its own stream), and a tombstone erases a key drawn uniformly from the
    keys inserted by earlier batches.  Lookups draw `hit_fraction` of their
    keys from the live keys and the rest uniformly from the absent keys of
    [0, 2^31).
    Range lower bounds are uniform, widths log-uniform in [wmin, wmax],
    hi = min(lo + w - 1, 2^31 - 1).
    """

    def __init__(self, seed, n_keys, batch, n_lookups, n_ranges, wmin, wmax,
                 erase_fraction=0.0, hit_fraction=0.5):
        self.seed = seed
        self.n_keys = n_keys
        self.batch = batch
        self.n_batches = (n_keys + batch - 1) // batch
        self.n_lookups = n_lookups
        self.n_ranges = n_ranges
        self.wmin, self.wmax = wmin, wmax
        self.erase_fraction = erase_fraction
        self.hit_fraction = hit_fraction

    def batches(self):
        """Yields (keys, values, is_erase) per batch."""
        pool = np.empty(self.n_keys, dtype=np.uint32)
        pooled = 0
        for t in range(self.n_batches):
            n = min(self.batch, self.n_keys - t * self.batch)
            base = t * self.batch
            keys = uniform31(self.seed, n, base)
            is_erase = np.zeros(n, dtype=np.uint8)
            if t > 0 and self.erase_fraction > 0 and pooled:
                r = stream(sub_seed(self.seed, 1), n, base)
                thresh = np.uint64(int(self.erase_fraction * 2 ** 32))
                sel = (r >> np.uint64(32)) < thresh
                pick = (r & np.uint64(0xFFFFFFFF)) % np.uint64(pooled)
                keys[sel] = pool[pick[sel].astype(np.int64)]
                is_erase[sel] = 1
            vals = murmur3_u32(keys)
            vals[is_erase == 1] = 0
            fresh = keys[is_erase == 0]
            pool[pooled:pooled + len(fresh)] = fresh
            pooled += len(fresh)
            yield keys, vals, is_erase

    def inserted_keys(self):
        """All keys inserted (by insert entries), in order."""
        parts = []
        for keys, _, e in self.batches():
            parts.append(keys[e == 0])
        return np.concatenate(parts) if parts else np.empty(0, np.uint32)

    @staticmethod
    def live_keys(batches):
        """Sorted keys live after replaying `batches` ((keys, values,
        is_erase) in order): the last operation on a key decides, so a key
        erased and inserted again is live."""
        if not batches:
            return np.empty(0, np.uint32)
        keys = np.concatenate([b[0] for b in batches])
        erase = np.concatenate([b[2] for b in batches])
        # last occurrence of each key: first occurrence in reversed order
        rk = keys[::-1]
        uniq, first = np.unique(rk, return_index=True)
        return uniq[erase[::-1][first] == 0]

    def lookups(self, live):
        """`hit_fraction` of the queries are drawn from `live` (sorted live
        keys, see live_keys); the rest are uniform draws from [0, 2^31)
        that are not live (present draws are redrawn from the stream's
        continuation), so the hit rate is exactly the fraction drawn."""
        n = self.n_lookups
        live = np.ascontiguousarray(live, dtype=np.uint32)
        r = stream(sub_seed(self.seed, 2), n)
        hit = (r >> np.uint64(32)) < np.uint64(int(self.hit_fraction * 2 ** 32))
        pick = (r & np.uint64(0xFFFFFFFF)) % np.uint64(max(len(live), 1))
        q = uniform31(sub_seed(self.seed, 3), n)
        if len(live):
            q[hit] = live[pick[hit].astype(np.int64)]
            miss = np.nonzero(~hit)[0]
            extra = n
            while len(miss):
                pos = np.searchsorted(live, q[miss])
                present = (pos < len(live)) & \
                    (live[np.minimum(pos, len(live) - 1)] == q[miss])
                miss = miss[present]
                q[miss] = uniform31(sub_seed(self.seed, 3), len(miss), extra)
                extra += len(miss)
        return q

    def ranges(self):
        n = self.n_ranges
        lo = uniform31(sub_seed(self.seed, 4), n)
        w = log_uniform(stream(sub_seed(self.seed, 5), n), self.wmin,
                        self.wmax)
        hi = np.minimum(lo.astype(np.uint64) + w - np.uint64(1),
                        np.uint64(2 ** 31 - 1)).astype(np.uint32)
        return lo, hi


def config1(scale_log=0):
    """2^20 keys in 16 batches of 2^16; 2^20 lookups (50% hit); 2^14 range
    counts with widths in [2^10, 2^16].  scale_log shrinks every size."""
    s = scale_log
    return LsmWorkload(SEEDS[1], 1 << (20 - s), 1 << (16 - s), 1 << (20 - s),
                       1 << (14 - s), 1 << 10, 1 << 16, 0.0, 0.5)


def config2(scale_log=0):
    """2^26 keys in batches of 2^20, 10% tombstones of earlier keys;
    2^26 lookups (70% live keys, 30% absent keys); 2^20 range counts with
    widths in [2^4, 2^20]."""
    s = scale_log
    return LsmWorkload(SEEDS[2], 1 << (26 - s), 1 << (20 - s), 1 << (26 - s),
                       1 << (20 - s), 1 << 4, 1 << 20, 0.1, 0.7)


class Zipf:
    """Zipf(s) ranks over 1..n by rejection-inversion (W. Hormann and
    G. Derflinger, ACM TOMACS 6(3), 1996).  Sample i draws its uniforms from
    the counter stream splitmix64(splitmix64(seed + i) + j), j = 0, 1, ...
    until one is accepted, so every sample is independent of the others
    and of the batch it is drawn in."""

    def __init__(self, s, n):
        self.s, self.n = float(s), float(n)
        self.h_x1 = self.H(1.5) - 1.0
        self.h_n = self.H(self.n + 0.5)
        self.sv = 2.0 - self.H_inv(self.H(2.5) - self.h(2.0))

    @staticmethod
    def _helper1(x):  # log1p(x) / x
        x = np.asarray(x, np.float64)
        small = np.abs(x) <= 1e-8
        safe = np.where(small, 1.0, x)
        return np.where(small, 1.0 - x * (0.5 - x / 3.0),
                        np.log1p(safe) / safe)

    @staticmethod
    def _helper2(x):  # expm1(x) / x
        x = np.asarray(x, np.float64)
        small = np.abs(x) <= 1e-8
        safe = np.where(small, 1.0, x)
        return np.where(small, 1.0 + x * 0.5 * (1.0 + x / 3.0),
                        np.expm1(safe) / safe)

    def h(self, x):
        return np.exp(-self.s * np.log(x))

    def H(self, x):
        lx = np.log(x)
        return self._helper2((1.0 - self.s) * lx) * lx

    def H_inv(self, x):
        t = np.maximum(x * (1.0 - self.s), -1.0)
        return np.exp(self._helper1(t) * x)

    def ranks(self, seed, count, offset=0):
        base = splitmix64(np.uint64(seed) +
                          np.arange(offset, offset + count, dtype=np.uint64))
        out = np.zeros(count, dtype=np.uint64)
        todo = np.arange(count)
        j = np.uint64(0)
        while len(todo):
            u01 = (splitmix64(base[todo] + j) >> np.uint64(11)).astype(
                np.float64) * 2.0 ** -53
            u = self.h_n + u01 * (self.h_x1 - self.h_n)
            x = self.H_inv(u)
            k = np.clip(np.floor(x + 0.5), 1.0, self.n)
            ok = (k - x <= self.sv) | (u >= self.H(k + 0.5) - self.h(k))
            out[todo[ok]] = k[ok].astype(np.uint64)
            todo = todo[~ok]
            j += np.uint64(1)
        return out


def _device_uniform01(torch, x):
    """u64 bit patterns (int64 tensor) -> uniforms in [0, 1): the top 53
    bits, as (x >> 11) * 2^-53 does for the numpy streams."""
    return ((x >> 11) & ((1 << 53) - 1)).to(torch.float64) * 2.0 ** -53


def zipf_ranks_device(zipf, torch, seed, count, dev, offset=0,
                      chunk=1 << 25):
    """Zipf.ranks on a GPU (torch, float64): the same counter streams and
    rejection-inversion steps, evaluated by the device's float64 log/exp
    (which may round differently from the host's in the last place, so a
    rare sample can land on a neighbouring rank).  For streams too large to
    draw on the host in seconds; returns an int64 tensor of ranks."""
    s = zipf.s

    def helper1(x):  # log1p(x) / x
        small = x.abs() <= 1e-8
        safe = torch.where(small, torch.ones_like(x), x)
        return torch.where(small, 1.0 - x * (0.5 - x / 3.0),
                           torch.log1p(safe) / safe)

    def helper2(x):  # expm1(x) / x
        small = x.abs() <= 1e-8
        safe = torch.where(small, torch.ones_like(x), x)
        return torch.where(small, 1.0 + x * 0.5 * (1.0 + x / 3.0),
                           torch.expm1(safe) / safe)

    def H(x):
        lx = torch.log(x)
        return helper2((1.0 - s) * lx) * lx

    def H_inv(x):
        t = torch.clamp(x * (1.0 - s), min=-1.0)
        return torch.exp(helper1(t) * x)

    def h(x):
        return torch.exp(-s * torch.log(x))

    out = torch.empty(count, dtype=torch.int64, device=dev)
    for a in range(0, count, chunk):
        n = min(chunk, count - a)
        idx = torch.arange(offset + a, offset + a + n, dtype=torch.int64,
                           device=dev)
        base = device_splitmix64(idx + _i64(int(seed) % (1 << 64)))
        res = torch.zeros(n, dtype=torch.int64, device=dev)
        todo = torch.arange(n, dtype=torch.int64, device=dev)
        j = 0
        while todo.numel():
            u01 = _device_uniform01(torch, device_splitmix64(base[todo] + j))
            u = zipf.h_n + u01 * (zipf.h_x1 - zipf.h_n)
            x = H_inv(u)
            k = torch.clamp(torch.floor(x + 0.5), 1.0, zipf.n)
            ok = (k - x <= zipf.sv) | (u >= H(k + 0.5) - h(k))
            res[todo[ok]] = k[ok].to(torch.int64)
            todo = todo[~ok]
            j += 1
        out[a:a + n] = res
    return out


def murmur3_u32_device(torch, k):
    """murmur3_u32 of uint32 keys held in an int64 tensor (values in
    [0, 2^32)); returns int64 values in [0, 2^32)."""
    m = (1 << 32) - 1

    def mul(a, c):
        return (a * c) & m

    def rotl(x, r):
        return ((x << r) | (x >> (32 - r))) & m
    k = mul(k, 0xCC9E2D51)
    k = rotl(k, 15)
    k = mul(k, 0x1B873593)
    h = k  # seed 0
    h = rotl(h, 13)
    h = (h * 5 + 0xE6546B64) & m
    h = h ^ 4
    h = h ^ (h >> 16)
    h = mul(h, 0x85EBCA6B)
    h = h ^ (h >> 13)
    h = mul(h, 0xC2B2AE35)
    return h ^ (h >> 16)


def scrambled_keys(ranks, bits, c=0x5BD1E995):
    """Rank r -> (r * 0x9E3779B1 + c) mod 2^bits: an odd multiplier makes
    it a bijection of the bits-wide key space (the paper-sweep key map,
    FliX/paper_lsmu_sweep.cu:64-67), so hot ranks land on scattered keys."""
    mask = np.uint64((1 << bits) - 1)
    with np.errstate(over="ignore"):
        return (np.asarray(ranks, np.uint64) * np.uint64(0x9E3779B1) +
                np.uint64(c)) & mask


class ZipfRequestStream:
    """The uneven request stream of config 4: `n` requests with sizes
    log-uniform in [smin, smax]; 40% inserts (values murmur3 of the key),
    10% erases, 40% lookups and 10% range counts (widths log-uniform in
    [2^4, 2^16]); every key is a Zipf(s) rank over 2^bits scrambled
    bijectively over the bits-wide key space."""

    def __init__(self, seed, n, s=0.99, bits=30, smin=32, smax=1 << 18):
        self.seed, self.n, self.bits = seed, n, bits
        self.smin, self.smax = smin, smax
        self.zipf = Zipf(s, 1 << bits)

    def requests(self):
        r = stream(sub_seed(self.seed, 30), self.n)
        sizes = log_uniform(stream(sub_seed(self.seed, 31), self.n),
                            self.smin, self.smax).astype(np.int64)
        u = (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        total = int(sizes.sum())
        keys = scrambled_keys(self.zipf.ranks(sub_seed(self.seed, 32), total),
                              self.bits).astype(np.uint32)
        widths = log_uniform(stream(sub_seed(self.seed, 33), total), 16,
                             1 << 16)
        out = []
        off = 0
        top = np.uint64((1 << 32) - 1)
        for i in range(self.n):
            n = int(sizes[i])
            k = keys[off:off + n]
            if u[i] < 0.4:
                out.append(("insert", k, murmur3_u32(k)))
            elif u[i] < 0.5:
                out.append(("erase", k, None))
            elif u[i] < 0.9:
                out.append(("lookup", k, None))
            else:
                hi = np.minimum(k.astype(np.uint64) + widths[off:off + n] -
                                np.uint64(1), top).astype(np.uint32)
                out.append(("range", k, hi))
            off += n
        return out


def _as_i32(torch, x):
    """[0, 2^32) values in an int64 tensor -> int32 with the same bits."""
    return (x - (x >= (1 << 31)).to(torch.int64) * (1 << 32)).to(torch.int32)


def requests_device(stream_obj, torch, dev):
    """ZipfRequestStream.requests generated on a GPU (the keys' Zipf ranks,
    the range widths and the insert values are drawn there; the request
    kinds and sizes on the host, identically): a list of (kind, a, b) with
    int32 device tensors holding the uint32 bits.  Same distributions and
    the same counter streams as requests(); a rare key can differ from the
    host draw by one Zipf rank (float64 rounding of the device's log/exp)."""
    w = stream_obj
    r = stream(sub_seed(w.seed, 30), w.n)
    sizes = log_uniform(stream(sub_seed(w.seed, 31), w.n), w.smin,
                        w.smax).astype(np.int64)
    u = (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    total = int(sizes.sum())
    ranks = zipf_ranks_device(w.zipf, torch, sub_seed(w.seed, 32), total, dev)
    mask = (1 << w.bits) - 1
    keys = (ranks * 0x9E3779B1 + 0x5BD1E995) & mask
    del ranks
    idx = torch.arange(total, dtype=torch.int64, device=dev)
    wu = _device_uniform01(torch, device_splitmix64(
        idx + _i64(sub_seed(w.seed, 33))))
    del idx
    widths = torch.clamp(torch.floor(torch.exp2(4.0 + wu * 12.0)), 16,
                         1 << 16).to(torch.int64)
    del wu
    out = []
    off = 0
    for i in range(w.n):
        n = int(sizes[i])
        k = keys[off:off + n]
        if u[i] < 0.4:
            out.append(("insert", _as_i32(torch, k),
                        _as_i32(torch, murmur3_u32_device(torch, k))))
        elif u[i] < 0.5:
            out.append(("erase", _as_i32(torch, k), None))
        elif u[i] < 0.9:
            out.append(("lookup", _as_i32(torch, k), None))
        else:
            hi = torch.clamp(k + widths[off:off + n] - 1, max=(1 << 32) - 1)
            out.append(("range", _as_i32(torch, k), _as_i32(torch, hi)))
        off += n
    return out


def config4(n_requests=10000, scale_log=0):
    """Config 4: 10^4 uneven requests (sizes log-uniform [32, 2^18]) over
    Zipf(0.99) keys on a 2^30 universe; scale_log shrinks the sizes."""
    return ZipfRequestStream(SEEDS[4], n_requests, 0.99, 30, 32,
                             max(32, (1 << 18) >> scale_log))


def _i64(c):
    """A uint64 constant as the int64 with the same bits."""
    return c - (1 << 64) if c >= (1 << 63) else c


def device_splitmix64(idx):
    """splitmix64 of an int64 torch tensor on its device, bit-identical to
    splitmix64() (int64 products wrap like uint64 ones; right shifts are
    masked to be logical)."""
    def lsr(x, k):
        return (x >> k) & ((1 << (64 - k)) - 1)
    x = idx + _i64(0x9E3779B97F4A7C15)
    x = (x ^ lsr(x, 30)) * _i64(0xBF58476D1CE4E5B9)
    x = (x ^ lsr(x, 27)) * _i64(0x94D049BB133111EB)
    return x ^ lsr(x, 31)


class ServingStream64:
    """Config 5 requests: sizes log-uniform in [1, 2^16]; 50% lookups, 30%
    inserts, 10% erases, 10% range sums; uint64 keys are Zipf(0.8) ranks r
    over `universe` mapped to splitmix64(r - 1), the keys an index grown
    with splitmix64(0 .. universe-1) holds, so hot ranks are hot stored
    keys; values splitmix64(key); range widths log-uniform in [2^4, 2^40]
    in key space."""

    def __init__(self, seed, n, universe, s=0.8, smin=1, smax=1 << 16):
        self.seed, self.n, self.universe = seed, n, universe
        self.smin, self.smax = smin, smax
        self.zipf = Zipf(s, universe)

    def requests(self):
        r = stream(sub_seed(self.seed, 40), self.n)
        sizes = log_uniform(stream(sub_seed(self.seed, 41), self.n),
                            self.smin, self.smax).astype(np.int64)
        u = (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        total = int(sizes.sum())
        ranks = self.zipf.ranks(sub_seed(self.seed, 42), total)
        keys = splitmix64(ranks - np.uint64(1))
        widths = log_uniform(stream(sub_seed(self.seed, 43), total), 16,
                             1 << 40)
        out = []
        off = 0
        top = np.uint64(0xFFFFFFFFFFFFFFFF)
        for i in range(self.n):
            n = int(sizes[i])
            k = keys[off:off + n]
            if u[i] < 0.5:
                out.append(("lookup", k, None))
            elif u[i] < 0.8:
                out.append(("insert", k, splitmix64(k)))
            elif u[i] < 0.9:
                out.append(("erase", k, None))
            else:
                hi = k + np.minimum(widths[off:off + n] - np.uint64(1),
                                    top - k)
                out.append(("range", k, hi))
            off += n
        return out


def config5(n_requests=2000, scale_log=0):
    """Config 5: a 2^31-key uint64 index served under Poisson arrivals."""
    return ServingStream64(SEEDS[5], n_requests, 1 << (31 - scale_log))


class RequestStream:
    """Mixed request stream for the serving comparison: `n` requests with
    sizes log-uniform in [smin, smax]; 40% inserts (new uniform 31-bit keys,
    values murmur3), 10% erases (keys drawn from the base keys and earlier
    inserts), 40% lookups (half drawn from known keys) and 10% range counts
    (widths log-uniform in [2^4, 2^16])."""

    KINDS = ("insert", "erase", "lookup", "range")

    def __init__(self, seed, n, smin=32, smax=1 << 18):
        self.seed, self.n, self.smin, self.smax = seed, n, smin, smax

    def requests(self, base_keys):
        r = stream(sub_seed(self.seed, 10), self.n)
        sizes = log_uniform(stream(sub_seed(self.seed, 11), self.n),
                            self.smin, self.smax).astype(np.int64)
        u = (r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
        base = np.ascontiguousarray(base_keys, dtype=np.uint32)
        B = len(base)
        ins = np.empty(int(sizes.sum()), dtype=np.uint32)  # inserted so far
        m = 0
        out = []
        off = 0
        for i in range(self.n):
            n = int(sizes[i])
            x = stream(sub_seed(self.seed, 12), 3 * n, off)
            off += 3 * n
            pick = (x[n:2 * n] % np.uint64(max(B + m, 1))).astype(np.int64)

            def known(p):
                from_ins = p >= B
                k = np.empty(len(p), dtype=np.uint32)
                if B:
                    k[~from_ins] = base[p[~from_ins]]
                k[from_ins] = ins[p[from_ins] - B]
                return k

            if u[i] < 0.4:
                k = (x[:n] >> np.uint64(33)).astype(np.uint32)
                out.append(("insert", k, murmur3_u32(k)))
                ins[m:m + n] = k
                m += n
            elif u[i] < 0.5:
                out.append(("erase", known(pick), None))
            elif u[i] < 0.9:
                k = (x[:n] >> np.uint64(33)).astype(np.uint32)
                half = n // 2
                k[:half] = known(pick[:half])
                out.append(("lookup", k, None))
            else:
                lo = (x[:n] >> np.uint64(33)).astype(np.uint32)
                w = log_uniform(x[2 * n:], 16, 1 << 16)
                hi = np.minimum(lo.astype(np.uint64) + w - np.uint64(1),
                                np.uint64(2 ** 31 - 1)).astype(np.uint32)
                out.append(("range", lo, hi))
        return out


class FlixWorkload:
    """Config 3 (FliX): a bulk build of `n_build` uint64 keys in clustered
    runs of 2^12 consecutive keys at random 64-bit offsets (kept inside
    [2, 2^64 - 4]), then batches of `batch` mixed ops (75% inserts of new
    uniform keys, 25% deletes of live built keys), each followed by
    `lookups` lookups of live keys and `ranges` range queries whose width
    targets 16..1024 pairs inside one run."""

    def __init__(self, seed, n_build, batch, lookups, ranges):
        self.seed, self.n_build, self.batch = seed, n_build, batch
        self.lookups, self.ranges = lookups, ranges

    # keys live in [2, 2^64 - 4]: 0 and 1 are reserved by FliX and the top
    # values are its sentinels (definitions.cuh:86-103)
    KEY_LO = 2
    KEY_SPAN = (1 << 64) - 4 - 2

    def build_keys(self):
        runs = max(1, self.n_build >> 12)
        st = stream(sub_seed(self.seed, 20), runs)
        # run starts uniform over the key range (a run of 4096 must fit);
        # overlapping runs are re-drawn from the next counter values
        span = np.uint64(self.KEY_SPAN - 4096)
        starts = st % span + np.uint64(self.KEY_LO)
        extra = 0
        while True:
            order = np.sort(starts)
            clash = np.nonzero(np.diff(order) < np.uint64(4096))[0]
            if len(clash) == 0:
                break
            bad = np.isin(starts, order[clash + 1])
            fresh = stream(sub_seed(self.seed, 23), int(bad.sum()), extra)
            extra += int(bad.sum())
            starts[bad] = fresh % span + np.uint64(self.KEY_LO)
        keys = (starts[:, None] + np.arange(4096, dtype=np.uint64)).ravel()
        return np.sort(keys[: self.n_build])

    def _built_order(self, j):
        """A bijection of [0, n) for n = len(built keys) = a power of two:
        j -> (j * 0x9E3779B1 + 0x7F4A7C15) mod n (an odd multiplier), the
        order in which built keys are deleted, so no key is deleted twice
        and the built keys at order positions >= the deletions so far are
        exactly the live ones."""
        n = np.uint64(self.n_build)
        with np.errstate(over="ignore"):
            return ((np.asarray(j, np.uint64) * np.uint64(0x9E3779B1) +
                     np.uint64(0x7F4A7C15)) % n).astype(np.int64)

    def deletes_per_batch(self):
        return self.batch - (3 * self.batch) // 4

    def batch_ops(self, t, built):
        """Batch t: 75% inserts of uniform keys over the key range (a
        uniform 64-bit key hits one of the built keys with probability
        < 2^-36, and an insert of a present key is a no-op in FliX and in
        the oracle alike), 25% deletes of live built keys, never deleted
        before (the next slice of the deletion order)."""
        n = self.batch
        n_ins = (3 * n) // 4
        nd = n - n_ins
        x = stream(sub_seed(self.seed, 21), n_ins, t * n_ins)
        ins = x % np.uint64(self.KEY_SPAN) + np.uint64(self.KEY_LO)
        dels = built[self._built_order(np.arange(t * nd, (t + 1) * nd))]
        vals = splitmix64(ins)
        return ins, vals, dels

    def queries(self, t, built):
        """After batch t: `lookups` lookups of live built keys and `ranges`
        range queries [lo, lo + w - 1] with lo a live built key and w
        log-uniform in [16, 1024] (inside a run of consecutive keys, so
        each returns up to w pairs)."""
        x = stream(sub_seed(self.seed, 22), self.lookups + 2 * self.ranges,
                   t * (self.lookups + 2 * self.ranges))
        gone = np.uint64(min((t + 1) * self.deletes_per_batch(),
                             self.n_build - 1))
        left = np.uint64(self.n_build) - gone

        def live(r):
            return built[self._built_order(gone + r % left)]

        q = live(x[: self.lookups])
        r = x[self.lookups:]
        lo = live(r[: self.ranges])
        w = log_uniform(r[self.ranges:], 16, 1024)
        room = np.uint64(self.KEY_LO + self.KEY_SPAN) - lo  # no wrap-around
        hi = lo + np.minimum(w - np.uint64(1), room)
        return q, lo, hi


def config3(scale_log=0):
    s = scale_log
    return FlixWorkload(SEEDS[3], 1 << (27 - s), 1 << (22 - s),
                        1 << (24 - s), 1 << (18 - s))
