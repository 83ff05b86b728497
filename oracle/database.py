"""Restatement of MEDYAN's element databases (src/Structure/Database.h). TEST INFRASTRUCTURE ONLY:
tests use it to derive the bead orders that serializeDof sees; the product never imports it.

* ElementList: DatabaseBase (Database.h:55-104). Creating an element appends it; destroying the
  element at index i moves the last element into i (when i is not last) and pops the end. This is
  the order of Bead::getBeads(), hence of serializeDof's coordinate vector (DofSerializer.hpp:78-84).
* StableIndex: the stable indices of Database<T, true> (Database.h:180-304). A created element
  fills the most recently deleted index (Database.h:192-207), else takes a new one at the end;
  a destroyed element only records its index (Database.h:296-299); rearrange() (Database.h:220-250)
  moves each surviving element at or above the final size, in ascending order, into the deleted
  indices below the final size, in the order they were recorded.
"""


class ElementList:
    def __init__(self):
        self.elems = []

    def create(self, e):
        self.elems.append(e)
        return len(self.elems) - 1

    def destroy(self, e):
        i = self.elems.index(e)
        if i + 1 != len(self.elems):
            self.elems[i] = self.elems[-1]
        self.elems.pop()

    def index(self, e):
        return self.elems.index(e)


class StableIndex:
    def __init__(self):
        self.stable = []     # element per stable index (None: deleted)
        self.deleted = []
        self.where = {}

    def create(self, e):
        if self.deleted:
            i = self.deleted.pop()
            self.stable[i] = e
        else:
            i = len(self.stable)
            self.stable.append(e)
        self.where[e] = i
        return i

    def destroy(self, e):
        i = self.where.pop(e)
        self.stable[i] = None
        self.deleted.append(i)

    def rearrange(self):
        final = len(self.stable) - len(self.deleted)
        gone = set(self.deleted)
        holes = [i for i in self.deleted if i < final]
        movers = [i for i in range(final, len(self.stable)) if i not in gone]
        assert len(holes) == len(movers)
        for h, m in zip(holes, movers):
            e = self.stable[m]
            self.stable[h] = e
            self.where[e] = h
        del self.stable[final:]
        self.deleted = []


def serialize_map(previous, following):
    """The mdn_dof_serialize map from the element order `previous` (the current system's beads) to
    `following` (the next system's): each element's index in `previous`, or -1 for a new one."""
    at = {e: i for i, e in enumerate(previous)}
    return [at.get(e, -1) for e in following]
