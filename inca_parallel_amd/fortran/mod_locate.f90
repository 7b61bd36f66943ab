!! Text search in formatted input files (behaviour of src/mod_locatemod.f90):
!! `locate` leaves the unit positioned after the first line containing the
!! key and stops the run if there is none; `located` only reports;
!! `getline_with` returns that line.
module locatemod
   implicit none
contains
   subroutine locate(iunit, key)
      integer, intent(in) :: iunit
      character(len=*), intent(in) :: key
      character(len=500) :: line
      if (.not. find_line(iunit, key, line)) then
         write(*,*) "ERROR: Section not found -> ", trim(key)
         stop 1
      end if
   end subroutine locate

   logical function located(iunit, key)
      integer, intent(in) :: iunit
      character(len=*), intent(in) :: key
      character(len=500) :: line
      located = find_line(iunit, key, line)
   end function located

   subroutine getline_with(iunit, key, line)
      integer, intent(in) :: iunit
      character(len=*), intent(in) :: key
      character(len=*), intent(out) :: line
      character(len=500) :: buf
      if (.not. find_line(iunit, key, buf)) then
         write(*,*) "ERROR: Section not found -> ", trim(key)
         stop 1
      end if
      line = buf
   end subroutine getline_with

   logical function find_line(iunit, key, line)
      integer, intent(in) :: iunit
      character(len=*), intent(in) :: key
      character(len=*), intent(out) :: line
      integer :: ios
      rewind(iunit)
      find_line = .false.
      do
         read(iunit, "(A)", iostat=ios) line
         if (ios /= 0) return
         if (index(line, key) /= 0) then
            find_line = .true.
            return
         end if
      end do
   end function find_line
end module locatemod
