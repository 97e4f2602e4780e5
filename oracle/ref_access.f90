! Read-only accessor compiled next to goSPL's fortran/functions.F90 inside
! oracle/_ref: copies the module's FV arrays (set by definetin) out to caller
! arrays so the test-suite can compare them with the C-ABI geometry.
subroutine ref_fv_export(nb, nnb, elgt, vdist, area)
  use meshparams
  implicit none
  integer :: nb
  integer, intent(out) :: nnb(nb)
  double precision, intent(out) :: elgt(nb,12), vdist(nb,12), area(nb)
  nnb = FVnNb
  elgt = FVeLgt
  vdist = FVvDist
  area = FVarea
end subroutine ref_fv_export
