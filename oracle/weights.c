/* Point weights: multiplicity, geometric factor h^2, cube-fraction cubature.
 * CPU restatement, test infrastructure only (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>

/* fixed_data.c:4244-4254: number of distinct orderings of (i,j,k) */
double or_permsix(int i, int j, int k)
{
    if (i == j && j == k) return 1.0;
    if (i == j || j == k || i == k) return 3.0;
    return 6.0;
}

/* calculate_geometric.c:59-92: the reference's large-l asymptotic form of
 * (2l1+1)(2l2+1)(2l3+1)/(4pi) (l1 l2 l3;0 0 0)^2; zero off the even-parity
 * triangle.  Kept as a reference-parity mode; the exact form is below. */
double or_geometric_asym(int l1, int l2, int l3)
{
    if ((l1 + l2 + l3) % 2 != 0 || l1 + l2 < l3 || l2 + l3 < l1 || l3 + l1 < l2) return 0.0;
    double L = (double)(l1 + l2 + l3);
    double a1 = (double)(l2 + l3 - l1), a2 = (double)(l3 + l1 - l2), a3 = (double)(l1 + l2 - l3);
    double pref = M_1_PI * M_1_PI / 2.0;
    double p1 = pref * ((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0)) / (L + 1.0);
    double p2 = (L + 1.0 / 3.0) / ((a1 + 1.0 / 3.0) * (a2 + 1.0 / 3.0) * (a3 + 1.0 / 3.0));
    double p3 = sqrt((a1 + 1.0 / 6.0) * (a2 + 1.0 / 6.0) * (a3 + 1.0 / 6.0) / (L + 1.0 / 6.0));
    return p1 * p2 * p3;
}

/* f(a) = C(2a,a)/4^a = prod_{k=1..a} (2k-1)/(2k).  With g = (l1+l2+l3)/2 and
 * a_i = g - l_i the textbook factorial formula for the squared 3j symbol
 * reduces to  (l1 l2 l3;0 0 0)^2 = f(a1) f(a2) f(a3) / ((2g+1) f(g)). */
static long double f_ld(int a)
{
    long double f = 1.0L;
    for (int k = 1; k <= a; ++k) f *= (long double)(2 * k - 1) / (long double)(2 * k);
    return f;
}

void or_threej_ftable(int amax, double* f)
{
    long double acc = 1.0L;
    f[0] = 1.0;
    for (int a = 1; a <= amax; ++a) {
        acc *= (long double)(2 * a - 1) / (long double)(2 * a);
        f[a] = (double)acc;
    }
}

double or_threej000_sq(int l1, int l2, int l3)
{
    int L = l1 + l2 + l3;
    if (L % 2 != 0 || l1 + l2 < l3 || l2 + l3 < l1 || l3 + l1 < l2) return 0.0;
    int g = L / 2;
    long double r = f_ld(g - l1) * f_ld(g - l2) * f_ld(g - l3) / ((long double)(L + 1) * f_ld(g));
    return (double)r;
}

double or_h2_exact(int l1, int l2, int l3)
{
    int L = l1 + l2 + l3;
    if (L % 2 != 0 || l1 + l2 < l3 || l2 + l3 < l1 || l3 + l1 < l2) return 0.0;
    int g = L / 2;
    long double r = f_ld(g - l1) * f_ld(g - l2) * f_ld(g - l3) / ((long double)(L + 1) * f_ld(g));
    long double pre = (long double)(2 * l1 + 1) * (long double)(2 * l2 + 1) * (long double)(2 * l3 + 1);
    return (double)(pre * r / (4.0L * 3.14159265358979323846264338327950288L));
}

/* Cell rules of calculate_volume.c:265-310 (weights of the 8 cube corners for
 * the fraction of a unit sub-cube inside the tetrapyd). */
static double cell1(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (11 * (a + b + c + e) + 17 * (d + g + f) + 25 * h) / 240.0; }
static double cell2(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (47 * a + 19 * (b + c + e) + 5 * (d + f + g) + h) / 720.0; }
static double cell3(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (40 * (b + c) + 35 * (a + d) + 26 * (f + g) + 19 * (e + h)) / 360.0; }
static double cell4(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (89 * a + 85 * (b + c + e) + 71 * (d + f + g) + 43 * h) / 720.0; }

/* calculate_volume.c:12-132.  Classify the 3x3x3 neighbourhood of (i,j,k):
 * -8 outside [min,max], 0 outside the triangle, 1 inside.  Each of the 8
 * sub-cubes sharing the centre contributes by its count of inside corners;
 * the corner values are the indicator of the centre (only pt[1][1][1] = 1),
 * stored as int as the reference does.  Times ordered multiplicity. */
double or_calculate_weight(int min, int max, int i, int j, int k)
{
    int grid[3][3][3];
    int inside = 0;
    for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) {
                int a = i + l - 1, b = j + m - 1, c = k + n - 1;
                if (a < min || a > max || b < min || b > max || c < min || c > max) grid[l][m][n] = -8;
                else if (a > b + c || b > a + c || c > a + b) grid[l][m][n] = 0;
                else { grid[l][m][n] = 1; ++inside; }
            }
    double w = 0.0;
    if (inside == 27) {
        w = 1.0;
    } else {
        int pt[3][3][3] = {{{0}}};
        pt[1][1][1] = 1;
        for (int s1 = 0; s1 < 2; ++s1)
            for (int s2 = 0; s2 < 2; ++s2)
                for (int s3 = 0; s3 < 2; ++s3) {
                    int sum = 0;
                    for (int l = 0; l < 2; ++l)
                        for (int m = 0; m < 2; ++m)
                            for (int n = 0; n < 2; ++n) sum += grid[l + s1][m + s2][n + s3];
                    int a = pt[s1][s2][s3], b = pt[s1 + 1][s2][s3], c = pt[s1][s2 + 1][s3];
                    int d = pt[s1 + 1][s2 + 1][s3], e = pt[s1][s2][s3 + 1];
                    int f = pt[s1 + 1][s2][s3 + 1], g = pt[s1][s2 + 1][s3 + 1];
                    int h = pt[s1 + 1][s2 + 1][s3 + 1];
                    if (sum == 4) {
                        if (grid[1 + s1][0 + s2][1 + s3] == 1) w += cell2(f, e, h, g, b, a, d, c);
                        if (grid[1 + s1][1 + s2][0 + s3] == 1) w += cell2(d, c, b, a, h, g, f, e);
                        if (grid[0 + s1][1 + s2][1 + s3] == 1) w += cell2(g, h, e, f, c, d, a, b);
                    } else if (sum == 5) {
                        w += cell1(a, b, c, d, e, f, g, h);
                    } else if (sum == 6) {
                        if (grid[0 + s1][0 + s2][1 + s3] == 1) w += cell3(g, h, e, f, c, d, a, b);
                        if (grid[1 + s1][0 + s2][0 + s3] == 1) w += cell3(f, h, b, d, e, g, a, c);
                        if (grid[0 + s1][1 + s2][0 + s3] == 1) w += cell3(d, h, c, g, b, f, a, e);
                    } else if (sum == 7) {
                        if (grid[0 + s1][1 + s2][0 + s3] == 0) w += cell4(f, e, h, g, b, a, d, c);
                        if (grid[0 + s1][0 + s2][1 + s3] == 0) w += cell4(d, c, b, a, h, g, f, e);
                        if (grid[1 + s1][0 + s2][0 + s3] == 0) w += cell4(g, h, e, f, c, d, a, b);
                    } else if (sum == 8) {
                        w += 1.0 / 8.0;
                    }
                }
    }
    if (i != k) w *= (i == j || j == k) ? 3.0 : 6.0;
    return w;
}
