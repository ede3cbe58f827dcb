* A small LP exercising every MPS feature the reader supports:
* OBJSENSE MAX, L/G/E rows, a free row, ranges on L and E rows,
* integrality markers, and the bound types UP, LO, MI, FR.
* Optimum (maximisation): x = (4, 0, 7, -2), c'x = -4, objective -4 + 10.
NAME          SMALLLP
OBJSENSE
    MAX
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
 E  R4
 N  FREE
COLUMNS
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    MARKER    'MARKER'     'INTORG'
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0
    MARKER    'MARKER'     'INTEND'
    X3        COST        -1.0   MYEQN        1.0
    X3        R4           1.0   FREE         3.0
    X4        COST         0.5   R4           2.0
RHS
    RHS       COST       -10.0
    RHS       LIM1         4.0   LIM2         1.0
    RHS       MYEQN        7.0   R4           3.0
RANGES
    RNG       LIM1         2.5   R4          -1.5
BOUNDS
 UP BND       X1           4.0
 LO BND       X2          -1.0
 UP BND       X2           1.0
 MI BND       X3
 UP BND       X3           8.0
 FR BND       X4
ENDATA
