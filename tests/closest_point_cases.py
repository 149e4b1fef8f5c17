"""Known answers of the reference's Triangle::closest_point (mesh.rs:45-117),
one per branch, worked out by hand in exact arithmetic.

Notation of mesh.rs:50-62: e0 = v1 - v0, e1 = v2 - v0, w = v0 - p,
a = e0.e0, b = e0.e1, c = e1.e1, d = e0.w, e = e1.w, det = ac - b^2,
s = be - cd, t = bd - ae.  The returned point is v0 + sc e0 + tc e1.

The reference clamps the unnormalised parameters of regions 2-6 to [0, a] or
[0, c] (mesh.rs:67,71,75,78,95,105) instead of [0, 1] after dividing by a or
c, so on triangles whose edges are not of unit length those branches return
points off the true closest point, some beyond the triangle.  The triangles
below have a, c != 1, and one of them an obtuse angle at the vertex that makes
the branch reachable (region 4a needs b < 0, 2a an obtuse angle at v2, 6a one
at v1).  Every point has z = 1/2 above the triangle's plane z = 0.

Each case: (name, triangle v0 v1 v2, point p, expected reference point,
true closest point).  The true closest point (Eberly's algorithm, which the
physics-fix mode uses) follows from the geometry and is given for the
branches where it differs.
"""
from __future__ import annotations

# A: v0=(1,1,0) v1=(3,1,0) v2=(3/2,4,0): e0=(2,0,0) e1=(1/2,3,0)
#    a=4, b=1, c=37/4, det=36
TRI_A = ((1.0, 1.0, 0.0), (3.0, 1.0, 0.0), (1.5, 4.0, 0.0))
# B: v0=(0,0,0) v1=(2,0,0) v2=(-1,3/2,0): e0=(2,0,0) e1=(-1,3/2,0)
#    a=4, b=-2, c=13/4, det=9 (obtuse at v0)
TRI_B = ((0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (-1.0, 1.5, 0.0))
# C: v0=(0,0,0) v1=(3,0,0) v2=(1,1/2,0): e0=(3,0,0) e1=(1,1/2,0)
#    a=9, b=3, c=5/4, det=9/4 (obtuse at v2)
TRI_C = ((0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (1.0, 0.5, 0.0))
# D: v0=(0,0,0) v1=(2,1/2,0) v2=(3,0,0): e0=(2,1/2,0) e1=(3,0,0)
#    a=17/4, b=6, c=9, det=9/4 (obtuse at v1)
TRI_D = ((0.0, 0.0, 0.0), (2.0, 0.5, 0.0), (3.0, 0.0, 0.0))

CASES = [
    # region 0 (s>=0, t>=0, s+t<det): p=(5/4,5/4,1/2): d=-1/2, e=-7/8,
    # s=15/4, t=3 -> sc=s/det=5/48, tc=t/det=1/12 -> the projection
    ("0", TRI_A, (1.25, 1.25, 0.5), (1.25, 1.25, 0.0), None),
    # region 1 (s>=0, t>=0, s+t>=det): p=(11/4,3/2,1/2): d=-7/2, e=-19/8,
    # s=30, t=6; numer=c+e-b-d=75/8, denom=a-2b+c=45/4 -> sc=5/6 (clamp
    # [0,1]), tc=1/6 -> v0 + 5/6 e0 + 1/6 e1 = (11/4, 3/2, 0)
    ("1", TRI_A, (2.75, 1.5, 0.5), (2.75, 1.5, 0.0), None),
    # region 2a (s<0, s+t>=det, c+e > b+d): tri C, p=(5/4,3/4,1/2): d=-15/4,
    # e=-13/8, s=-3/16, t=27/8; tmp0=b+d=-3/4, tmp1=c+e=-3/8;
    # sc=(tmp1-tmp0)/(a-2b+c)=(3/8)/(17/4)=3/34, tc=31/34 -> (20/17, 31/68, 0)
    ("2a", TRI_C, (1.25, 0.75, 0.5), (20.0 / 17.0, 31.0 / 68.0, 0.0), None),
    # region 2b (s<0, s+t>=det, c+e <= b+d): p=(5/4,9/2,1/2): d=-1/2,
    # e=-85/8, s=-6, t=42; tmp0=1/2, tmp1=-11/8 -> tc=clamp(85/8, 0, c=37/4)
    # =37/4, sc=0 -> v0 + 37/4 e1 = (45/8, 115/4, 0); true closest: v2
    ("2b", TRI_A, (1.25, 4.5, 0.5), (45.0 / 8.0, 115.0 / 4.0, 0.0), (1.5, 4.0, 0.0)),
    # region 3 (s<0, t>=0, s+t<det): p=(0,7/4,1/2): d=2, e=-7/4, s=-81/4,
    # t=9 -> sc=0, tc=clamp(7/4, 0, 37/4)=7/4 -> v0 + 7/4 e1 = (15/8, 25/4, 0);
    # true closest: on edge v0v2 at parameter 7/37 -> (1 + 7/74, 1 + 21/37, 0)
    ("3", TRI_A, (0.0, 1.75, 0.5), (15.0 / 8.0, 25.0 / 4.0, 0.0), (1.0 + 7.0 / 74.0, 1.0 + 21.0 / 37.0, 0.0)),
    # region 4a (s<0, t<0, d<0): tri B, p=(3/4,-5/4,1/2): d=-3/2, e=21/8,
    # s=-3/8, t=-15/2 -> sc=clamp(3/2, 0, a=4)=3/2, tc=0 -> (3, 0, 0), beyond
    # v1=(2,0,0); true closest: on edge v0v1 at parameter 3/8 -> (3/4, 0, 0)
    ("4a", TRI_B, (0.75, -1.25, 0.5), (3.0, 0.0, 0.0), (0.75, 0.0, 0.0)),
    # region 4b (s<0, t<0, d>=0): tri B, p=(-3/2,-1/4,1/2): d=3, e=-9/8,
    # s=-15/2, t=-3/2 -> sc=0, tc=clamp(9/8, 0, 13/4)=9/8 -> 9/8 e1 =
    # (-9/8, 27/16, 0); true closest: on edge v0v2 at parameter -e/c = 9/26
    # -> (-9/26, 27/52, 0)
    ("4b", TRI_B, (-1.5, -0.25, 0.5), (-9.0 / 8.0, 27.0 / 16.0, 0.0), (-9.0 / 26.0, 27.0 / 52.0, 0.0)),
    # region 5 (s>=0, t<0, s+t<det): p=(7/4,0,1/2): d=-3/2, e=21/8, s=33/2,
    # t=-12 -> sc=clamp(3/2, 0, 4)=3/2, tc=0 -> v0 + 3/2 e0 = (4, 1, 0), beyond
    # v1=(3,1,0); true closest: on edge v0v1 at 3/8 -> (7/4, 1, 0)
    ("5", TRI_A, (1.75, 0.0, 0.5), (4.0, 1.0, 0.0), (1.75, 1.0, 0.0)),
    # region 6a (s>=0, t<0, s+t>=det, a+d > b+e): tri D, p=(9/4,3/4,1/2):
    # d=-39/8, e=-27/4, s=27/8, t=-9/16; numer=c+e-b-d=9/8, denom=a-2b+c=5/4
    # -> sc=9/10, tc=1/10 -> (21/10, 9/20, 0)
    ("6a", TRI_D, (2.25, 0.75, 0.5), (2.1, 0.45, 0.0), None),
    # region 6b (s>=0, t<0, s+t>=det, a+d <= b+e): p=(19/4,3/4,1/2):
    # d=-15/2, e=-9/8, s=273/4, t=-3; a+d=-7/2 <= b+e=-1/8 -> sc=clamp(9/8,
    # 0, c=37/4)=9/8 (clamp(-e) on the e0 parameter, mesh.rs:105), tc=0 ->
    # v0 + 9/8 e0 = (13/4, 1, 0), beyond v1=(3,1,0); true closest: v1
    ("6b", TRI_A, (4.75, 0.75, 0.5), (3.25, 1.0, 0.0), (3.0, 1.0, 0.0)),
]


def tri12(tri, normal=(0.0, 0.0, 1.0)):
    """Triangle as the 12 numbers (v0, v1, v2, normal) of the oracle's rpo_triangle."""
    return [*tri[0], *tri[1], *tri[2], *normal]
