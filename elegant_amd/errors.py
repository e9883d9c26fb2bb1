"""Seeded element errors in the manner of elegant's error_element command.

* Uniform stream: random_1_elegant (src/drand_oag.c:94-117) over LAPACK's
  DLARAN (src/drand_oag.c:119-191), a multiplicative congruential generator
  x <- 33952834046453 x mod 2^48 with 12-bit limbs.  Reproduced exactly for the
  seed 987654321, the one seed the reference uses without permuting its bits
  (src/drand_oag.c:81-82, :103-104); other seeds go through the SDDS toolkit's
  permuteSeedBitOrder, which is not in the reference snapshot, so they are taken
  as given (bit permutation skipped) and flagged.
* Gaussian errors: gauss_rn_lim of the SDDS toolkit (not in the snapshot;
  parity unpinned).  Restated here as the polar Box-Muller method on the same
  uniform stream, one value kept for the next call, redrawing values beyond the
  cutoff.
* perturbation() (src/error.c:496-534): uniform 2A(u - 1/2), Gaussian with
  cutoff, plus-or-minus A.
* assert_perturbations() (src/vary.c:797-960): per error item, occurrences of
  each matching element name in beamline order; with bind (the default) one
  value per name (every bind_number occurrences if bind_number >= 1);
  fractional errors multiply the nominal value by 1 + delta, additive ones add
  delta, non-additive ones replace the value.
"""
import fnmatch
import math

from . import structs as S

UNPERMUTED_SEED = 987654321


class Random1:
    """random_1_elegant: 48-bit state in four 12-bit limbs, DLARAN step"""

    def __init__(self, seed=UNPERMUTED_SEED):
        seed = abs(int(seed))
        # True when the stream is the reference's own (no SDDS bit permutation involved)
        self.pinned = seed == UNPERMUTED_SEED
        seed = (seed // 2) * 2 + 1
        self.seed = [(seed >> 36) & 4095, (seed >> 24) & 4095, (seed >> 12) & 4095, seed & 4095]

    def __call__(self):
        i1, i2, i3, i4 = self.seed
        it4 = i4 * 2549
        it3 = it4 // 4096
        it4 -= it3 << 12
        it3 = it3 + i3 * 2549 + i4 * 2508
        it2 = it3 // 4096
        it3 -= it2 << 12
        it2 = it2 + i2 * 2549 + i3 * 2508 + i4 * 322
        it1 = it2 // 4096
        it2 -= it1 << 12
        it1 = it1 + i1 * 2549 + i2 * 2508 + i3 * 322 + i4 * 494
        it1 %= 4096
        self.seed = [it1, it2, it3, it4]
        return (it1 + (it2 + (it3 + it4 * 2.44140625e-4) * 2.44140625e-4) * 2.44140625e-4) * 2.44140625e-4


class Gauss:
    """Gaussian deviates from a uniform stream (polar method), cut at `limit` sigmas"""

    def __init__(self, urandom):
        self.u = urandom
        self.saved = None

    def unit(self):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            v1 = 2 * self.u() - 1
            v2 = 2 * self.u() - 1
            r = v1 * v1 + v2 * v2
            if 0 < r < 1:
                break
        fac = math.sqrt(-2 * math.log(r) / r)
        self.saved = v1 * fac
        return v2 * fac

    def lim(self, mean, sigma, limit):
        while True:
            v = self.unit()
            if limit <= 0 or abs(v) <= limit:
                return mean + sigma * v


class ErrorItem:
    """one error_element command: name may use wildcards; item is the parameter
    (elegant's name, e.g. FSE, DX, K1)"""

    def __init__(self, name, item, amplitude, cutoff=3.0, type="gaussian", bind=1, bind_number=0, fractional=False,
                 additive=True, element_type=None):
        if type not in ("gaussian", "uniform", "plus_or_minus"):
            raise ValueError("error type %r not supported" % type)
        if fractional and not additive:
            raise ValueError("fractional errors must be additive")
        self.name, self.item, self.amplitude, self.cutoff = name, item.upper(), float(amplitude), float(cutoff)
        self.type, self.bind, self.bind_number = type, int(bind), int(bind_number)
        self.fractional, self.additive, self.element_type = bool(fractional), bool(additive), element_type


class ErrorSet:
    def __init__(self, items, seed=UNPERMUTED_SEED):
        self.items = list(items)
        self.rng = Random1(seed)
        self.gauss = Gauss(self.rng)

    def perturbation(self, it):
        if it.type == "uniform":
            return 2 * it.amplitude * (self.rng() - 0.5)
        if it.type == "gaussian":
            return self.gauss.lim(0.0, it.amplitude, it.cutoff)
        return it.amplitude * (1.0 if self.rng() > 0.5 else -1.0)

    @staticmethod
    def _value(el, item):
        st = S.make_struct(el.type, el.params)
        alias = S.ALIASES.get(el.type, {})
        field = alias.get(item) or ("length" if item == "L" else item.lower())
        return getattr(st, field)

    def apply(self, lattice):
        """perturb the lattice in place; returns [(element index, item, delta, new value)]"""
        log = []
        for it in self.items:
            names = []
            for e in lattice.elements:
                if fnmatch.fnmatchcase(e.name, it.name) and e.name not in names and (
                        it.element_type is None or S.TYPE_NAMES[e.type] == it.element_type.upper()):
                    names.append(e.name)
            for nm in names:
                i_group = 0
                delta = 0.0
                for i, e in enumerate(lattice.elements):
                    if e.name != nm:
                        continue
                    if (not it.bind) or (it.bind_number >= 1 and i_group % it.bind_number == 0) or i_group == 0:
                        delta = self.perturbation(it)
                    value = self._value(e, it.item)
                    if it.fractional:
                        value *= 1 + delta
                    elif not it.additive:
                        value = delta
                    else:
                        value += delta
                    lattice.set_param(i, it.item, value)
                    log.append((i, it.item, delta, value))
                    i_group += 1
                    if it.bind == -1:
                        delta = -delta
        return log
