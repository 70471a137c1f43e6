# kprep.awk -- the two source transforms that the reference applies to every generated kernel before
# compiling it (reference Makefile:157-162), restated:
#  1. in the double half (before "#define REAL QREAL"), an accumulation into a |.| companion array
#     ("; n_<node>_Wa[..]..+=") is made conditional on the thread-local SQ_AB flag;
#  2. the QREAL half is compiled at the optimisation level QOPT (a GCC pragma before it);
#  3. the graph_t descriptor is declared extern, so that it has external linkage in C++.
/^#define REAL QREAL/ { inq = 1; if (QOPT != "") printf "#pragma GCC optimize (\"%s\")\n", QOPT }
{
    line = $0
    if (!inq && index(line, "_Wa[") > 0) {
        out = ""
        while (match(line, /; n_[A-Za-z0-9_]+_Wa(\[[^]]*\])+\+=/)) {
            out = out substr(line, 1, RSTART + 1) "if(SQ_AB) " substr(line, RSTART + 2, RLENGTH - 2)
            line = substr(line, RSTART + RLENGTH)
        }
        line = out line
    }
    sub(/^const graph_t G_/, "extern const graph_t G_", line)
    print line
}
