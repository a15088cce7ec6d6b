"""River discharge at the mouths from the reference's river table
(src/fv_rivers.F90), for the host to hand to the step as the per-node field
"runoff" (key_rivers).

MESHPATH/rivers.dat, read list-directed (each read starts on a new line and
may run over several lines; rivers_ini, src/fv_rivers.F90:75-214):

    Ntime_rivers                          number of records
    Nnods_rivers                          number of river mouths
    time_rivers(1:Ntime)                  record times, julian days
    nodes_rivers(1:Nnods)                 mouth nodes, 1-based
    Ntime reads of Nnods discharges       m3/s, one record each

The discharge starts at zero (rivers_ini, 210-212).  After each step's clock
advance (src/fv_main.F90:400-412) rivers_do (263-300) interpolates linearly
between the records that bracket time_jd (getcoeffld, 216-261, with the
record search of binarysearch, 305-350), scaled by (time_jd - time_jd0)/10
during the first ten days.  Past the last record, or with a single record,
getcoeffld sets one_field and rivers_do stops updating: the discharge keeps
its last value (zero if that happens at start-up, as in the reference).
Model times before the first record have no record to start from (the
reference reads out of bounds there): this module raises ValueError.

The step takes one value per node.  Rivers that share a mouth are summed
here; the reference adds them one by one, which can differ in the last bit.
"""
import numpy as np

# binarysearch's tolerance, a REAL(4) literal (src/fv_rivers.F90:330)
_TOL = float(np.float32(1e-9))


def _list_directed(lines):
    """Reader of Fortran list-directed records: read(n) starts a new line and
    takes n values (separators blank or comma, r*c repeats)."""
    it = iter(lines)

    def read(n, conv):
        vals = []
        while len(vals) < n:
            line = next(it, None)
            if line is None:
                raise ValueError("rivers.dat: file ends inside a read")
            for tok in line.replace(",", " ").split():
                if tok.startswith("/"):
                    break
                if "*" in tok:
                    r, c = tok.split("*", 1)
                    vals.extend([conv(c.replace("d", "e").replace("D", "e"))] * int(r))
                else:
                    vals.append(conv(tok.replace("d", "e").replace("D", "e")))
        return vals[:n]
    return read


def binarysearch(array, value):
    """1-based index of the record at time value (within the tolerance), else
    of the last record before it, 0 if none (src/fv_rivers.F90:305-350)."""
    left, right = 1, len(array)
    while left <= right:
        middle = (left + right + 1) // 2     # nint((left+right)/2.0)
        if abs(array[middle - 1] - value) <= _TOL:
            return middle
        if array[middle - 1] > value:
            right = middle - 1
        else:
            left = middle + 1
    return right


class Rivers:
    """rivers_ini / rivers_do of src/fv_rivers.F90 for a serial host."""

    def __init__(self, path, nod2D, time_jd, time_jd0):
        read = _list_directed(open(path).read().splitlines())
        self.ntime = read(1, int)[0]
        nnods = read(1, int)[0]
        self.time = np.array(read(self.ntime, float))
        self.nodes = np.array(read(nnods, int)) - 1
        self.data = np.array([read(nnods, float) for _ in range(self.ntime)]).reshape(self.ntime, nnods)
        if nnods and (self.nodes.min() < 0 or self.nodes.max() >= nod2D):
            raise ValueError("rivers.dat: mouth node outside 1..nod2D")
        self.nod2D, self.time_jd0 = nod2D, float(time_jd0)
        self.one_field = False
        self.runoff_rivers = np.zeros(nnods)
        if nnods:
            self._coeff(float(time_jd))

    def _coeff(self, rdate):
        """getcoeffld (src/fv_rivers.F90:216-261)."""
        t = binarysearch(self.time, rdate)
        if t == 0:
            raise ValueError(f"model time {rdate} before the first river record {self.time[0]}")
        if t < self.ntime:
            p1 = t + 1
            delta = self.time[p1 - 1] - self.time[t - 1]
        else:
            p1, delta = t, 1.0
            self.one_field = True
        if rdate < self.time[0]:
            p1, delta = t, 1.0
        self.t_indx, self.t_indx_p1 = t, p1
        self.coef_a = (self.data[p1 - 1] - self.data[t - 1]) / delta
        self.coef_b = self.data[t - 1] - self.coef_a * self.time[t - 1]

    def do(self, time_jd):
        """rivers_do (src/fv_rivers.F90:263-300) at the advanced clock; returns
        the per-node discharge [m3/s]."""
        rdate = float(time_jd)
        if len(self.nodes) and not self.one_field:
            if rdate > self.time[self.t_indx_p1 - 1]:
                self._coeff(rdate)
            r = rdate * self.coef_a + self.coef_b
            if rdate - self.time_jd0 < 10.0:
                r = r * (rdate - self.time_jd0) / 10.0
            self.runoff_rivers = r
        return self.per_node()

    def per_node(self):
        out = np.zeros(self.nod2D)
        np.add.at(out, self.nodes, self.runoff_rivers)
        return out
