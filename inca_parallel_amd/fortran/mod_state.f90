!! Global state of the INCA host, with the module and variable names of the
!! reference (src/mod_geninfo.f90, mod_wfxinfo.f90, mod_intrainfo.f90,
!! mod_inputdat.f90, mod_quadratures.f90, mod_numbers.f90) so that code written
!! against the reference's globals keeps working.  The grid arrays that the
!! reference keeps in `quadratures` now live on the GPU; the host copies are
!! filled only where an output writer needs them.

module numbers
   implicit none
   double precision, parameter :: ZERO = 0.0d0, ONE = 1.0d0, TWO = 2.0d0, HALF = 0.5d0
   double precision, parameter :: pi = 4.d0*datan(1.d0)
contains
   !! Double factorial n!! with (-1)!! = 0!! = 1 (src/mod_numbers.f90:14-27).
   integer function dfact(n)
      integer, intent(in) :: n
      integer :: k
      dfact = 1
      k = n
      do while (k > 1)
         dfact = dfact*k
         k = k - 2
      end do
   end function dfact
end module numbers

module geninfo
   implicit none
   logical :: uhf = .false., rhf = .false., rdens = .false., udens = .false.
   logical :: opsh = .false., clsh = .false.
   integer :: natoms = 0, nelec = 0, nprim = 0, nalfae = 0, nbetae = 0
   double precision :: netch = 0.d0
   integer, allocatable :: Ra(:), Ptyp(:), an(:)
   integer, allocatable :: TMN(:,:)          ! (nprim,3)
   double precision, allocatable :: Alpha(:), cartes(:,:), Xn(:), Yn(:), Zn(:), chrg(:)
end module geninfo

module wfxinfo
   implicit none
   logical :: corr = .false.
   integer :: noccmo = 0
   double precision :: toteng = 0.d0
   double precision, allocatable :: T(:,:), Occ(:)
end module wfxinfo

module intrainfo
   implicit none
   character(len=40) :: dm2name = '', outname = ''
   character(len=32) :: calc_type = ''
   double precision :: thresh = 1.0d-8
   logical :: nohydro = .true., betaone = .false.
   logical :: radial_integral = .false., radial_plot = .false., cubeintra = .false.
   logical :: intracule_at_zero = .false., intracule_two_points = .false.
   integer :: nquad = 0
   double precision, allocatable :: cent(:,:), sfalpha(:), Ps(:)
   integer, allocatable :: nradc(:), nangc(:)
   integer :: nrad = 50, nang = 590
   character(len=40) :: r_plot_name = ''
   integer :: nblock = 1
   integer, allocatable :: n_an_per_part(:)
   double precision, allocatable :: tart(:,:), stp(:)
   character(len=40) :: cubeintraname = ''
   double precision :: center_i(3) = 0.d0, step_i(3) = 0.d0
   integer :: np_i(3) = 0
   double precision :: x_point1 = 0.d0, y_point1 = 0.d0, z_point1 = 0.d0
   double precision :: x_point2 = 0.d0, y_point2 = 0.d0, z_point2 = 0.d0
   ! engine options of this build (no reference counterpart)
   logical :: mixed_precision = .false.
   double precision :: mixed_tail = 0.02d0
   logical :: exact_polynomials = .false.
   double precision :: point_cutoff = 100.0d0   ! log2 of the per-point cutoff; 0 = off
   logical :: engine_grouping = .true.          ! evaluate by shell-quartet groups
   character(len=256) :: learned_grid_weights = ''  ! .npy of the learned autogrid network
   logical :: learned_grid_used = .false.
end module intrainfo

module inputdat
   implicit none
   logical :: readwfx = .false., readfchk = .false., readlog = .false., readbas = .false.
   logical :: cube = .false., primcube = .false., aocube = .false., mocube = .false.
   logical :: denscube = .false., gradient = .false., laplacian = .false.
   logical :: intracalc = .false., c1calc = .false.
   character(len=40) :: wfxfilename = '', fchkfilename = '', logfilename = '', basfilename = ''
end module inputdat

module quadratures
   implicit none
   integer :: maxgrid = 0, rgrid = 0, rrgrid = 0
   logical :: nosym = .false., definite = .false.
   integer :: nradi = 0
   double precision, allocatable :: radi(:)
end module quadratures
