"""Bootstrap draws of p_hat from the choices behind it.

Contract (reference: solver/inference/draws.py): the two table shapes, the weight schemes, the
messages for bad tables, and, above all, the random streams and the float64 values of every draw,
which the reference's stored records pin through sha256 fingerprints of the B x M draw matrix.

    panel    columns respondent, menu, chosen: respondents are resampled as units (Draws)
    replies  columns menu, chosen: each menu's replies are resampled on their own (MenuDraws)

Layout here. A table is parsed once into a ChoiceTable (codes, menus, chosen, checked). Its link to
rum's rows is a _RowMap: for every table row, the index of its menu among rum's menus and the rum row
of its (menu, chosen) cell, or -1. Panel draws are two sparse count matrices (respondents x menus
offered, respondents x rum rows chosen) multiplied by the B x R weights, a block of draws at a time;
with integer weights every entry is an exact integer, so the one division per entry fixes the value.
Replies draws are, per block k of draws, one numpy Generator seeded [seed, k] that draws every menu's
counts in increasing mask order.
"""
import hashlib

import numpy as np
import scipy.sparse as sp

WEIGHTS = ("multinomial_reject", "multinomial", "exponential", "poisson", "mammen")
COLUMNS = ("respondent", "menu", "chosen")
P_HAT_TOL = 1e-9                # rum's p_hat against the counts of the table
BLOCK_BYTES = 1 << 27           # a block of draws is at most 128 MB (it also fixes MenuDraws' streams)
DENSE_ENTRIES = 1 << 25         # count matrices up to this many entries are kept dense
REDRAW_ROUNDS = 10000           # multinomial_reject stops redrawing after this many rounds


def _first(mask):
    """Index of the first True entry, or None."""
    hit = np.flatnonzero(mask)
    return int(hit[0]) if hit.size else None


def _as_int64(x, column, what):
    """An integer column as int64 (integral floats allowed), with the table's messages."""
    kind = x.dtype.kind
    if kind == "f":
        i = _first(~np.isfinite(x))
        if i is not None:
            raise ValueError("%s row %d: missing %s" % (what, i, column))
        i = _first(x != np.floor(x))
        if i is not None:
            raise ValueError("%s row %d: %s %r is not an integer" % (what, i, column, x[i]))
    elif kind not in "iu":
        raise ValueError("the %s's %s column must hold integers" % (what, column))
    return x.astype(np.int64)


class ChoiceTable:
    """One checked table of choices. menu and chosen are int64; for a panel, codes are the
    respondents numbered 0..R-1 in sorted id order."""

    def __init__(self, table, what, num_alternatives=None, panel=False):
        names = COLUMNS if panel else COLUMNS[1:]
        absent = [c for c in names if c not in table]
        if absent:
            raise ValueError("the %s needs the columns %s and %s; missing: %s"
                             % (what, ", ".join(names[:-1]), names[-1], ", ".join(absent)))
        raw = {c: np.asarray(table[c]) for c in names}
        if any(v.ndim != 1 for v in raw.values()) or len({v.size for v in raw.values()}) != 1:
            raise ValueError("the %s's columns must be 1-D and of equal length" % what)
        if raw["menu"].size == 0:
            raise ValueError("the %s has no rows" % what)
        menu = _as_int64(raw["menu"], "menu", what)
        chosen = _as_int64(raw["chosen"], "chosen", what)
        # each check runs only when the ones before it passed (the later ones assume them)
        checks = (
            (lambda: (menu <= 0) | (menu >= 1 << 62),
             lambda i: "%s row %d: menu mask %d is not a nonempty set of at most 62 alternatives" % (what, i, menu[i])),
            (lambda: (chosen < 0) | (chosen > 61),
             lambda i: "%s row %d: chosen %d is not an alternative (0 to 61)" % (what, i, chosen[i])),
            (lambda: ((menu >> chosen) & 1) == 0,
             lambda i: "%s row %d: chosen %d is not in its menu (mask %d)" % (what, i, chosen[i], menu[i])),
        )
        if num_alternatives is not None:
            checks += ((lambda: menu >= 1 << int(num_alternatives),
                        lambda i: "%s row %d: menu mask %d has alternatives beyond num_alternatives = %d"
                        % (what, i, menu[i], num_alternatives)),)
        for bad, message in checks:
            i = _first(bad())
            if i is not None:
                raise ValueError(message(i))
        self.menu, self.chosen, self.codes, self.R = menu, chosen, None, None
        if panel:
            self.codes, self.R = self._respondents(raw["respondent"])

    @staticmethod
    def _respondents(resp):
        if resp.dtype.kind == "f" and not np.isfinite(resp).all():
            raise ValueError("panel row %d: missing respondent" % _first(~np.isfinite(resp)))
        if resp.dtype.kind == "O":
            for i, r in enumerate(resp):
                if r is None or r != r:
                    raise ValueError("panel row %d: missing respondent" % i)
        try:
            ids, codes = np.unique(resp, return_inverse=True)
        except TypeError:
            raise ValueError("respondent ids must be all numbers or all strings") from None
        return codes.astype(np.int64), int(ids.size)

    def fingerprint(self):
        """sha256[:16] over the codes (panel), menus and choices as stored: the records' input key."""
        parts = ([self.codes.tobytes()] if self.codes is not None else []) + [self.menu.tobytes(),
                                                                             self.chosen.tobytes()]
        return hashlib.sha256(b"".join(parts)).hexdigest()[:16]


def read_panel(panel, num_alternatives=None):
    """The panel's columns, checked: (respondent codes 0..R-1 in sorted id order, R, menu, chosen)."""
    t = ChoiceTable(panel, "panel", num_alternatives, panel=True)
    return t.codes, t.R, t.menu, t.chosen


def cells_from_panel(table, num_alternatives=None, alternatives=None):
    """(menu, alternative, p_hat) counted from a panel or a replies table: one row per menu of the
    table and member alternative (or only those in `alternatives`), menus in increasing mask order,
    alternatives increasing within a menu; p_hat = the share of the menu's rows choosing it."""
    panel = COLUMNS[0] in table
    t = ChoiceTable(table, "panel" if panel else "replies table", num_alternatives, panel=panel)
    n = (int(num_alternatives) if num_alternatives is not None
         else max(int(t.menu.max()).bit_length(), int(t.chosen.max()) + 1))
    keep = np.arange(n) if alternatives is None else np.unique(np.asarray(alternatives, dtype=np.int64))
    if keep.size == 0 or keep[0] < 0 or keep[-1] >= n:
        raise ValueError("alternatives must be a nonempty list of alternatives from 0 to %d" % (n - 1))
    menus, rows_of_menu = np.unique(t.menu, return_counts=True)
    offered = ((menus[:, None] >> keep[None, :]) & 1).astype(bool)       # menu x kept alternative
    if not offered.any():
        raise ValueError("no menu of the table offers any of the alternatives %s" % keep.tolist())
    # counts of each (menu, alternative) choice, keyed menu index * 64 + alternative
    key, count = np.unique(np.searchsorted(menus, t.menu) * 64 + t.chosen, return_counts=True)
    mi, ki = np.nonzero(offered)
    want = mi * 64 + keep[ki]
    at = np.minimum(np.searchsorted(key, want), key.size - 1)
    num = np.where(key[at] == want, count[at], 0)
    return menus[mi], keep[ki], num / rows_of_menu[mi]


class _RowMap:
    """How a table's rows meet rum's rows: menu_of[i] = index of table row i's menu among rum's
    menus (valid where on_menu[i]); cell_of[i] = rum row of table row i's (menu, chosen), or -1."""

    def __init__(self, table, menu, alternative):
        self.menus, self.row_menu = np.unique(menu, return_inverse=True)
        idx = np.minimum(np.searchsorted(self.menus, table.menu), self.menus.size - 1)
        self.on_menu = self.menus[idx] == table.menu
        self.menu_of = idx
        cell_key = self.row_menu * 64 + alternative
        by_key = np.argsort(cell_key)
        sorted_keys = cell_key[by_key]
        want = idx * 64 + table.chosen
        at = np.minimum(np.searchsorted(sorted_keys, want), sorted_keys.size - 1)
        self.cell_of = np.where(self.on_menu & (sorted_keys[at] == want), by_key[at], -1)

    def check_p_hat(self, menu, alternative, p_hat, counts, totals, what):
        """rum's p_hat must be the table's shares (within P_HAT_TOL)."""
        gap = np.abs(counts / totals - p_hat)
        if gap.max() > P_HAT_TOL:
            k = int(np.argmax(gap))
            source = "the panel gives" if what == "panel" else "the replies give"
            raise ValueError("rum row %d (menu %d, alternative %d): p_hat = %.6g, but %s %.6g (%d of %d rows); "
                             "build rum with cells_from_panel(%s, ...)"
                             % (k, menu[k], alternative[k], p_hat[k], source, counts[k] / totals[k], counts[k],
                                totals[k], what))


class _Blocks:
    """Draws built a block of rows_per_block draws at a time by _build(k) -> (block, draws in it
    that left a menu unseen). A full pass of blocks() sets fingerprint: sha256[:16] of the B x M
    float64 draws in row order."""

    rows_per_block = n_blocks = 1
    fingerprint = n_empty = None
    _kept = (None, None, 0)

    def _plan_blocks(self, width):
        self.rows_per_block = max(1, min(self.B, BLOCK_BYTES // (8 * width)))
        self.n_blocks = -(-self.B // self.rows_per_block)

    def block(self, k):
        if self._kept[0] != k:
            self._kept = (k,) + tuple(self._build(k))
        return self._kept[1]

    def blocks(self):
        digest, empty = hashlib.sha256(), 0
        for k in range(self.n_blocks):
            out = self.block(k)
            empty += self._kept[2]
            digest.update(np.ascontiguousarray(out).tobytes())
            lo = k * self.rows_per_block
            yield lo, lo + out.shape[0], out
        self.fingerprint, self.n_empty = digest.hexdigest()[:16], empty

    def rows(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        out = np.empty((idx.size, self.M))
        which = idx // self.rows_per_block
        for k in np.unique(which):
            sel = which == k
            out[sel] = self.block(int(k))[idx[sel] - k * self.rows_per_block]
        return out


def _multinomial_rows(rng, R, count):
    return rng.multinomial(R, np.ones(R) / R, size=count).astype(np.float64)


def _weights_mammen(rng, B, R, _unseen):
    r5 = np.sqrt(5.0)
    return np.where(rng.random((B, R)) < (r5 + 1) / (2 * r5), -(r5 - 1) / 2, (r5 + 1) / 2), 0, 0


def _weights_exponential(rng, B, R, _unseen):
    return rng.exponential(1.0, size=(B, R)), 0, 0


def _weights_poisson(rng, B, R, _unseen):
    return rng.poisson(1.0, size=(B, R)).astype(np.float64), 0, 0


def _weights_multinomial(rng, B, R, _unseen):
    return _multinomial_rows(rng, R, B), 0, 0


def _weights_multinomial_reject(rng, B, R, unseen):
    """Multinomial resamples; the rows that leave a menu of rum unseen are drawn again, all of
    them at once per round, until none is left (at most REDRAW_ROUNDS rounds)."""
    W = _multinomial_rows(rng, R, B)
    redrawn = rounds = 0
    while rounds < REDRAW_ROUNDS:
        bad = unseen(W)
        nbad = int(bad.sum())
        if nbad == 0:
            break
        W[bad] = _multinomial_rows(rng, R, nbad)
        redrawn, rounds = redrawn + nbad, rounds + 1
    return W, redrawn, rounds


_SCHEMES = {"mammen": _weights_mammen, "exponential": _weights_exponential, "poisson": _weights_poisson,
            "multinomial": _weights_multinomial, "multinomial_reject": _weights_multinomial_reject}


def _maybe_dense(a, entries):
    return a.toarray() if entries <= DENSE_ENTRIES else a


class Draws(_Blocks):
    """B draws from a panel: respondent i gets weight W[b, i] and draw b is
    p_hat_b(D, a) = sum_i W_bi chose_i(D, a) / sum_i W_bi offered_i(D)  (p_hat where the sum is 0);
    mammen: p_hat_b = clip(p_hat + sum_i W_bi (chose_i - offered_i p_hat) / N_D, 0, 1)."""

    input = "panel"

    def __init__(self, panel, menu, alternative, p_hat, B=10000, weights="multinomial_reject", seed=0):
        if weights not in WEIGHTS:
            raise ValueError("unknown weights %r; choose from %s" % (weights, WEIGHTS))
        if int(B) != B or B < 1:
            raise ValueError("B must be a positive integer")
        t = ChoiceTable(panel, "panel", panel=True)
        self.input_fingerprint = t.fingerprint()
        menu, alternative = np.asarray(menu, dtype=np.int64), np.asarray(alternative, dtype=np.int64)
        self.p_hat = p_hat = np.asarray(p_hat, dtype=np.float64)
        self.B, self.R, self.M, self.weights, self.seed = int(B), t.R, menu.size, weights, seed
        rm = _RowMap(t, menu, alternative)
        R, M, nm = t.R, self.M, rm.menus.size
        on, hit = rm.on_menu, rm.cell_of >= 0
        offered = sp.csr_array((np.ones(int(on.sum())), (t.codes[on], rm.menu_of[on])), shape=(R, nm))
        chose = sp.csr_array((np.ones(int(hit.sum())), (t.codes[hit], rm.cell_of[hit])), shape=(R, M))
        N = np.asarray(offered.sum(axis=0)).ravel()[rm.row_menu]
        k = _first(N == 0)
        if k is not None:
            raise ValueError("rum row %d (menu %d): the menu is not in the panel" % (k, menu[k]))
        rm.check_p_hat(menu, alternative, p_hat, np.asarray(chose.sum(axis=0)).ravel(), N, "panel")
        self._row_menu = rm.row_menu
        self._offered = _maybe_dense(offered, R * nm)
        self._seen = None
        if weights in ("mammen", "exponential"):
            # offered counts per rum row: offered x (menu -> its rows)
            spread = sp.csr_array((np.ones(M), (rm.row_menu, np.arange(M))), shape=(nm, M))
            seen = offered @ spread
        if weights == "mammen":
            self._resid = _maybe_dense((chose - seen.multiply(p_hat[None, :])).tocsr(), R * M)
            self._N = np.maximum(N, 1.0)
        else:
            self._chose = _maybe_dense(chose, R * M)
            if weights == "exponential":
                self._seen = _maybe_dense(seen, R * M)
        self.rows_unused = int((~on).sum())
        self._plan_blocks(M)
        self.W, self.n_redraws, self.n_reject_rounds = _SCHEMES[weights](np.random.default_rng(seed), self.B, R,
                                                                         self._unseen)

    def _unseen(self, W):
        """Rows of W that give some menu of rum a total weight <= 0."""
        step = max(1, BLOCK_BYTES // (8 * self._offered.shape[1]))
        return np.concatenate([np.any(W[a:a + step] @ self._offered <= 0, axis=1) for a in range(0, W.shape[0], step)])

    def _build(self, k):
        Wb = self.W[k * self.rows_per_block:(k + 1) * self.rows_per_block]
        if self.weights == "mammen":
            return np.clip(self.p_hat + np.asarray(Wb @ self._resid) / self._N, 0.0, 1.0), 0
        total = (np.asarray(Wb @ self._offered)[:, self._row_menu] if self._seen is None
                 else np.asarray(Wb @ self._seen))
        chosen = np.asarray(Wb @ self._chose)
        empty = total <= 0
        return np.where(empty, self.p_hat, chosen / np.where(empty, 1.0, total)), int(empty.any(axis=1).sum())

    @property
    def default_n(self):
        return self.R

    def check_n(self, n):
        if self.R != n:
            raise ValueError("n = %d, but the panel has %d respondents" % (n, self.R))

    def summary(self):
        return dict(respondents=self.R, rows_unused=self.rows_unused, redraws=self.n_redraws,
                    emptied_menu_draws=self.n_empty)

    @property
    def describe(self):
        return "%s draws over respondents" % self.weights

    def __repr__(self):
        fp = "" if self.fingerprint is None else ", fingerprint " + self.fingerprint
        return "Draws(B=%d, %s, seed %s: %d respondents, %d rows, %d per block%s)" % (
            self.B, self.weights, self.seed, self.R, self.M, self.rows_per_block, fp)


class MenuDraws(_Blocks):
    """B draws from replies: in draw b, menu D's counts are Multinomial(N_D, shares of D's replies
    over the alternatives chosen there), independently across menus and draws; block k draws every
    menu in increasing mask order from numpy.random.default_rng([seed, k])."""

    input, weights = "replies", "multinomial"
    n_redraws = 0

    def __init__(self, replies, menu, alternative, p_hat, B=10000, seed=0):
        if int(B) != B or B < 1:
            raise ValueError("B must be a positive integer")
        t = ChoiceTable(replies, "replies table")
        self.input_fingerprint = t.fingerprint()
        menu, alternative = np.asarray(menu, dtype=np.int64), np.asarray(alternative, dtype=np.int64)
        self.p_hat = p_hat = np.asarray(p_hat, dtype=np.float64)
        self.B, self.M, self.seed = int(B), menu.size, seed
        rm = _RowMap(t, menu, alternative)
        on = rm.on_menu
        self.N = np.bincount(rm.menu_of[on], minlength=rm.menus.size)      # replies per rum menu
        k = _first(self.N[rm.row_menu] == 0)
        if k is not None:
            raise ValueError("rum row %d (menu %d): the menu is not in the replies" % (k, menu[k]))
        # the (menu, alternative) pairs chosen at least once, menu-major: the multinomial categories
        cat, count = np.unique(rm.menu_of[on] * 64 + t.chosen[on], return_counts=True)
        cat_menu = cat // 64
        want = rm.row_menu * 64 + alternative
        at = np.minimum(np.searchsorted(cat, want), cat.size - 1)
        found = cat[at] == want
        totals = self.N[rm.row_menu]
        rm.check_p_hat(menu, alternative, p_hat, np.where(found, count[at], 0), totals, "replies")
        self._shares = count / self.N[cat_menu]
        self._span = np.r_[np.searchsorted(cat_menu, np.arange(self.N.size)), cat.size]
        self._cat_of_row = np.where(found, at, -1)
        self._totals = totals.astype(np.float64)
        self.rows_unused = int((~on).sum())
        self._plan_blocks(max(self.M, cat.size))

    def _build(self, k):
        rows = min(self.rows_per_block, self.B - k * self.rows_per_block)
        rng = np.random.default_rng([self.seed, k])
        counts = np.empty((rows, self._shares.size))
        for j, (a, b) in enumerate(zip(self._span[:-1], self._span[1:])):
            counts[:, a:b] = rng.multinomial(self.N[j], self._shares[a:b], size=rows)
        picked = counts[:, np.maximum(self._cat_of_row, 0)]
        return np.where(self._cat_of_row >= 0, picked, 0.0) / self._totals, 0

    @property
    def default_n(self):
        return int(self.N.max())

    def check_n(self, n):
        if not self.N.min() <= n <= self.N.max():
            raise ValueError("n = %d, but the menus have %d to %d replies; with replies, n is the number of replies "
                             "per menu" % (n, self.N.min(), self.N.max()))

    def summary(self):
        return dict(replies=int(self.N.sum()), replies_per_menu=[int(self.N.min()), int(self.N.max())],
                    rows_unused=self.rows_unused)

    @property
    def describe(self):
        return "draws menu by menu"

    def __repr__(self):
        fp = "" if self.fingerprint is None else ", fingerprint " + self.fingerprint
        return "MenuDraws(B=%d, seed %s: %d replies on %d menus, %d to %d each; %d rows, %d per block%s)" % (
            self.B, self.seed, self.N.sum(), self.N.size, self.N.min(), self.N.max(), self.M, self.rows_per_block, fp)
