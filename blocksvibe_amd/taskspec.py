"""blocks_3d's command-line contract: one invocation's options, read from argv and an optional `-p` param file,
normalised into a TaskSpec.

The contract itself -- option names, boost-style unique-prefix matching, `key = value` param files whose values
only fill options the command line left unset, the error messages and the output-file template -- is
blocks_3d's and the reference's (reference blocksvibe/cli.py:29-207, blocks_3d's handle_input), so pipelines
built on it keep working.  Here it is read in two passes: `read_options` turns argv (then the param file) into a
name -> text map through the option table OPTIONS; TaskSpec converts and checks the values.
"""
import re
from typing import NamedTuple


class UsageError(ValueError):
    """a command line or param file blocks_3d would reject (message as blocks_3d's)"""


class Opt(NamedTuple):
    name: str
    default: object = None          # None: required
    short: str = ""
    flag: bool = False              # takes no value


REQUIRED = None
OPTIONS = (
    Opt("j-external"), Opt("j-internal"), Opt("j-12"), Opt("j-43"), Opt("delta-12"), Opt("delta-43"),
    Opt("delta-1-plus-2", "0"), Opt("four-pt-struct"), Opt("four-pt-sign"), Opt("order"), Opt("lambda"),
    Opt("coordinates", "zzb"), Opt("kept-pole-order"), Opt("num-threads"), Opt("precision"),
    Opt("out-template", short="o"), Opt("debug", "false"), Opt("profile", "true"),
    Opt("param-file", "", short="p"), Opt("help", "", short="h", flag=True),
)
BY_NAME = {o.name: o for o in OPTIONS}
BY_SHORT = {o.short: o for o in OPTIONS if o.short}
ALL_COORDS = ("zzb", "yyb", "xt", "ws", "xt_radial", "ws_radial")
GPU_COORDS = frozenset(("xt", "xt_radial"))


def resolve(name):
    """an option's full name from an exact name or a unique prefix"""
    if name in BY_NAME:
        return name
    hits = [n for n in BY_NAME if n.startswith(name)]
    if len(hits) == 1:
        return hits[0]
    raise UsageError(f"unrecognised option '--{name}'")


def _pairs(argv):
    """(option name, value text or None for a flag) for each option of argv"""
    it = iter(argv)
    for word in it:
        if word.startswith("--"):
            name, has_eq, text = word[2:].partition("=")
            opt = BY_NAME[resolve(name)]
            value = text if has_eq else None
        elif len(word) >= 2 and word[0] == "-" and word[1] in BY_SHORT:
            opt = BY_SHORT[word[1]]
            value = word[2:] or None
        else:
            raise UsageError(f"unexpected argument '{word}'")
        if opt.flag:
            yield opt.name, None
            continue
        if value is None:
            value = next(it, None)
            if value is None:
                raise UsageError(f"option '--{opt.name}' requires an argument")
        yield opt.name, value


def read_param_file(path):
    """`key = value` lines; '#' starts a comment, blank lines and `[section]` lines are skipped"""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        raise UsageError(f"Could not open '{path}'")
    out = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or line.startswith("["):
            continue
        key, has_eq, value = line.partition("=")
        key = key.strip()
        if not has_eq or key not in BY_NAME:
            raise UsageError(f"unrecognised line in param file: '{line}'")
        out.setdefault(key, value.strip())
    return out


def read_options(argv):
    """name -> value text: the command line, then the param file for the options it did not set"""
    given = {}
    for name, value in _pairs(argv):
        if name == "help":
            given["help"] = "1"
            continue
        if name in given:
            raise UsageError(f"option '--{name}' cannot be specified more than once")
        given[name] = value
    if given.get("param-file"):
        for key, value in read_param_file(given["param-file"]).items():
            given.setdefault(key, value)
    return given


def doubled(text):
    """2 x a whole or half integer given as text"""
    x = 2 * float(text)
    if x != int(x):
        raise UsageError(f"not a whole or half integer: {text}")
    return int(x)


def boolean(text):
    word = text.strip().lower()
    if word in ("1", "true", "yes", "on"):
        return True
    if word in ("0", "false", "no", "off"):
        return False
    raise UsageError(f"not a boolean: {text}")


def spin_set(spec, two_j12):
    """the doubled internal spins of a j-internal list: items 'a', 'a-b' or 'a-b-step', comma separated"""
    spins = set()
    for item in spec.split(","):
        fields = item.split("-")
        if not item or len(fields) > 3:
            raise UsageError(f"Found invalid element in j-internal: '{item}'")
        lo = doubled(fields[0])
        hi = doubled(fields[1]) if len(fields) >= 2 else lo
        if lo % 2 != two_j12 % 2:
            raise UsageError("j_12 and j_internal must all be either whole or half integers")
        step = 2 * int(fields[2]) if len(fields) == 3 else 2
        if (hi - lo) % step:
            raise UsageError(f"The step of a spin range in j_internal must exactly divide the difference: {item}")
        spins.update(range(lo, hi + 1, step))
    return sorted(spins)


_FIELD = re.compile(r"\{(?::([^}]*))?\}")


def fill_template(template, x):
    """fmt::format(template, x) for a double x and the '{}' / '{:spec}' fields blocks_3d's templates use"""
    def field(m):
        if m.group(1):
            return format(x, m.group(1))
        text = repr(float(x))
        return text[:-2] if text.endswith(".0") else text
    return _FIELD.sub(field, template)


class TaskSpec:
    """one blocks_3d invocation: spins doubled (two_js, two_j12, two_j43, q4, two_js_internal), Delta values kept
    as the text given (they are printed back verbatim), sign, orders, coordinates, precision, template"""

    def __init__(self, argv, prog="blocks_3d"):
        self.argv = [prog] + list(argv)
        opts = read_options(argv)
        self.help = "help" in opts
        if self.help:
            return
        for o in OPTIONS:
            if o.default is REQUIRED and o.name not in opts:
                raise UsageError(f"the option '--{o.name}' is required but missing")
        val = lambda name: opts.get(name, BY_NAME[name].default)
        self.two_j12, self.two_j43 = doubled(val("j-12")), doubled(val("j-43"))
        if (self.two_j12 - self.two_j43) % 2:
            raise UsageError("j_12 and j_43 must both be either integers or half integers")
        self.two_js = [doubled(x) for x in val("j-external").split(",")]
        self.q4 = [doubled(x) for x in val("four-pt-struct").split(",")]
        if len(self.two_js) != 4 or len(self.q4) != 4:
            raise UsageError("j-external and four-pt-struct take 4 values")
        self.sign = int(val("four-pt-sign"))
        if self.sign not in (1, -1):
            raise UsageError(f"Four-pt-sign must be 1 or -1, but got: {self.sign}")
        for i, (q, j) in enumerate(zip(self.q4, self.two_js)):
            if (q - j) % 2 or abs(q) > j:
                raise UsageError(f"invalid four-pt-struct entry {i}")
        self.two_js_internal = spin_set(val("j-internal"), self.two_j12)
        self.delta_12, self.delta_43 = val("delta-12"), val("delta-43")
        self.delta_1_plus_2 = val("delta-1-plus-2")
        self.order, self.Lambda = int(val("order")), int(val("lambda"))
        self.kept_pole_order, self.num_threads = int(val("kept-pole-order")), int(val("num-threads"))
        self.precision = int(val("precision"))
        self.out_template = val("out-template")
        self.coordinates = set(val("coordinates").split(","))
        unknown = sorted(self.coordinates - set(ALL_COORDS))
        if unknown:
            raise UsageError(f"Invalid entry for coordinates: '{unknown[0]}'")
        self.radial = all(c.endswith("_radial") for c in self.coordinates)
        self.debug, self.profile = boolean(val("debug")), boolean(val("profile"))
        self.p0 = (1 - self.sign) // 2
        self.tjmax = self.two_js_internal[-1]

    def unsupported(self):
        """why the GPU path cannot run this task (None: it can)"""
        other = sorted(self.coordinates - GPU_COORDS)
        return f"coordinates {other} (GPU: xt, xt_radial)" if other else None

    def group_key(self):
        """tasks with this key share the recursion operator (M matrices, plan, constants)"""
        return (tuple(self.two_js), self.two_j12, self.two_j43, self.delta_12, self.delta_43, self.order,
                self.Lambda, self.tjmax, self.precision, min(self.kept_pole_order, self.order))

    def lambda_orders(self):
        return [self.p0] if self.radial else list(range(self.p0, self.Lambda + 1, 2))

    def path(self, x):
        return fill_template(self.out_template, x)

    def coords_in_order(self):
        return [c for c in ALL_COORDS if c in self.coordinates]
