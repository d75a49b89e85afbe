"""Alignments, data parts and the Data container (p4/alignment.py,
p4/part.py, p4/data.py), restricted to what the likelihood path needs.

An Alignment holds equal-length sequences of one data type.  Each
Alignment (or each site block of it) becomes one Part; Part.cPart is built
exactly as Alignment._initParts does (p4/alignment.py:5784-5880): newPart,
pokeEquatesTable, pokeSequences, makePatterns, setGlobalInvarSitesVec.
"""
import numpy

from . import pf
from .p4exceptions import P4Error

DNA_EQUATES = {'n': 'acgt', 'm': 'ac', 'k': 'gt', 'h': 'act', 'y': 'ct', 'v': 'acg',
               'w': 'at', 'd': 'agt', 'b': 'cgt', 'r': 'ag', 's': 'cg'}
PROTEIN_EQUATES = {'b': 'dn', 'x': 'arndcqeghilkmfpstwyv', 'z': 'eq'}
DNA_SYMBOLS = 'acgt'
PROTEIN_SYMBOLS = 'arndcqeghilkmfpstwyv'


class Alignment(object):
    """Equal-length sequences.  dataType 'dna', 'protein' or 'standard'
    (then give symbols).  Sequences are lower-cased as p4 does."""

    def __init__(self, names, sequences, dataType='dna', symbols=None, equates=None):
        if len(names) != len(sequences):
            raise P4Error("Alignment: names and sequences differ in number")
        self.taxNames = list(names)
        self.sequences = [s.lower() for s in sequences]
        lens = set(len(s) for s in self.sequences)
        if len(lens) != 1:
            raise P4Error("Alignment: sequences differ in length")
        self.length = lens.pop()
        self.dataType = dataType
        if dataType == 'dna':
            self.symbols = DNA_SYMBOLS
            self.equates = dict(DNA_EQUATES) if equates is None else equates
        elif dataType == 'protein':
            self.symbols = PROTEIN_SYMBOLS
            self.equates = dict(PROTEIN_EQUATES) if equates is None else equates
        elif dataType == 'standard':
            if not symbols:
                raise P4Error("Alignment: 'standard' data needs symbols")
            self.symbols = symbols
            self.equates = dict(equates or {})
        else:
            raise P4Error("Alignment: unknown dataType %s" % dataType)
        self.dim = len(self.symbols)

    @property
    def nTax(self):
        return len(self.sequences)

    @classmethod
    def fromCodes(cls, names, codes, dataType='dna', symbols=None):
        """Build from an int array [nTax][nSites] of state indices
        (0..dim-1); -1 is a gap, -2 a '?'."""
        sym = {'dna': DNA_SYMBOLS, 'protein': PROTEIN_SYMBOLS}.get(dataType, symbols)
        lut = numpy.array(list(sym) + ['?', '-'])
        codes = numpy.asarray(codes)
        idx = numpy.where(codes >= 0, codes, numpy.where(codes == -1, len(sym) + 1, len(sym)))
        seqs = [''.join(lut[row]) for row in idx]
        return cls(names, seqs, dataType, symbols)

    def _makePart(self, start=0, end=None):
        end = self.length if end is None else end
        p = Part()
        p.dim = self.dim
        p.symbols = self.symbols
        p.equates = self.equates
        p.dataType = self.dataType
        p.nTax = self.nTax
        p.nChar = end - start
        eqSymb = list(self.equates.keys())
        p.cPart = pf.newPart(self.nTax, p.nChar, ''.join(eqSymb), len(eqSymb), self.symbols, self.dim)
        table = []
        for e in eqSymb:
            allowed = self.equates[e]
            for s in self.symbols:
                table.append('1' if s in allowed else '0')
        pf.pokeEquatesTable(p.cPart, ''.join(table))
        pf.pokeSequences(p.cPart, ''.join(s[start:end] for s in self.sequences))
        pf.makePatterns(p.cPart)
        pf.setGlobalInvarSitesVec(p.cPart)
        p.nPatterns = pf.partPatternCount(p.cPart)
        return p


class Part(object):
    def __init__(self):
        self.cPart = None
        self.dim = 0
        self.nTax = 0
        self.nChar = 0
        self.nPatterns = 0
        self.symbols = None
        self.equates = None
        self.dataType = None

    def __del__(self):
        if getattr(self, "cPart", None):
            try:
                pf.freePart(self.cPart)
            except Exception:
                pass
            self.cPart = None

    def patterns(self):
        """(patterns, counts, seqPosPatternIndex, invarVec, invarArray)."""
        return pf.partGetPatterns(self.cPart, self.nTax)

    @classmethod
    def fromPatterns(cls, patterns, counts, dataType='dna', symbols=None, equates=None):
        """A part from already-compressed coded patterns [nTax][nPatterns]
        (codes as Pf/defines.h:33-36) and their counts."""
        patterns = numpy.ascontiguousarray(patterns, dtype=numpy.int32)
        counts = numpy.ascontiguousarray(counts, dtype=numpy.int32)
        sym = {'dna': DNA_SYMBOLS, 'protein': PROTEIN_SYMBOLS}.get(dataType, symbols)
        eq = {'dna': DNA_EQUATES, 'protein': PROTEIN_EQUATES}.get(dataType, equates or {})
        p = cls()
        p.dim = len(sym)
        p.symbols = sym
        p.equates = eq
        p.dataType = dataType
        p.nTax = patterns.shape[0]
        p.nChar = int(counts.sum())
        eqSymb = list(eq.keys())
        p.cPart = pf.newPart(p.nTax, p.nChar, ''.join(eqSymb), len(eqSymb), sym, p.dim)
        table = []
        for e in eqSymb:
            for s in sym:
                table.append('1' if s in eq[e] else '0')
        pf.pokeEquatesTable(p.cPart, ''.join(table))
        pf.partSetPatterns(p.cPart, patterns, counts)
        pf.setGlobalInvarSitesVec(p.cPart)
        p.nPatterns = patterns.shape[1]
        return p


class Data(object):
    """Container of parts over a common taxon set (p4/data.py)."""

    def __init__(self, alignments=None, parts=None, taxNames=None):
        self.parts = []
        self.alignments = []
        self.cData = None
        if alignments is not None:
            if isinstance(alignments, Alignment):
                alignments = [alignments]
            self.alignments = list(alignments)
            names = alignments[0].taxNames
            for a in alignments:
                if a.taxNames != names:
                    raise P4Error("Data: alignments have different taxa")
                self.parts.append(a._makePart())
            self.taxNames = list(names)
        elif parts is not None:
            self.parts = list(parts)
            self.taxNames = list(taxNames) if taxNames else ["t%i" % i for i in range(parts[0].nTax)]
        else:
            raise P4Error("Data needs alignments or parts")
        nTax = set(p.nTax for p in self.parts)
        if len(nTax) != 1:
            raise P4Error("Data: parts have different numbers of taxa")
        self.nTax = nTax.pop()

    @property
    def nParts(self):
        return len(self.parts)

    def dupe(self):
        """A copy with new C parts (p4/data.py:670-683)."""
        if self.alignments:
            als = [Alignment(a.taxNames, list(a.sequences), a.dataType, a.symbols, dict(a.equates))
                   for a in self.alignments]
            return Data(alignments=als)
        parts = []
        for p in self.parts:
            pats, cnts, _, _, _ = p.patterns()
            parts.append(Part.fromPatterns(pats, cnts, p.dataType, p.symbols, p.equates))
        return Data(parts=parts, taxNames=self.taxNames)

    def resetSequencesFromParts(self):
        """The alignments' sequences from their parts' C sequences (after a
        simulation; p4/data.py:305-310, p4/alignment.py)."""
        for a, p in zip(self.alignments, self.parts):
            s = pf.symbolSequences(p.cPart)
            n = pf.lib().p4g_partNChar(p.cPart)
            a.sequences = [s[i * n:(i + 1) * n] for i in range(p.nTax)]
            a.length = n

    def simpleBigXSquared(self):
        """Per part, the compositional chi-square over taxa, no gaps or
        ambiguities allowed (p4/data.py:636-651; -2.0 if there are)."""
        return [pf.partBigXSquared(p.cPart) for p in self.parts]

    def simpleConstantSitesCount(self):
        """Per part, sites whose non-gap characters are all one state
        (p4/data.py:653-668)."""
        return [pf.partSimpleConstantSitesCount(p.cPart) for p in self.parts]

    def meanNCharsPerSite(self):
        """Per part, the mean number of different states per site
        (p4/data.py:752-767)."""
        return [pf.partMeanNCharsPerSite(p.cPart) for p in self.parts]

    def _setCStuff(self):
        """p4/data.py:261-278."""
        self.cData = pf.newData(self.nTax, self.nParts)
        for i, p in enumerate(self.parts):
            pf.pokePartInData(p.cPart, self.cData, i)

    def __del__(self):
        if getattr(self, "cData", None):
            try:
                pf.freeData(self.cData)
            except Exception:
                pass
            self.cData = None
